"""CPU checks of the C ABI: the library loads without a GPU, exports every
function the headers declare, and its host-only input utilities (scalar
generator, reduction mod l) match Python integers and the oracle."""
import os
import re

import numpy as np
import pytest

import edref
import libftcrypto_amd as ft
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("ftcrypto.h", "ftcrypto_ext.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(ftx?_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_reference_surface():
    names = declared()
    ref38 = ["ft_error_str", "ft_crypto_init", "ft_crypto_device_memory", "ft_array_init_scalar",
             "ft_array_init_point", "ft_array_get_subset", "ft_array_get_range", "ft_array_assign_to_subset",
             "ft_array_assign_zero_to_subset", "ft_array_assign_to_range", "ft_array_delete_subset",
             "ft_array_delete_range", "ft_array_to_bytes_folded", "ft_array_to_bytes_affine",
             "ft_array_to_bytes", "ft_array_from_bytes_folded", "ft_array_from_bytes_affine",
             "ft_array_from_bytes", "ft_array_validate", "ft_array_from_scalars",
             "ft_array_from_small_scalars", "ft_array_free", "ft_array_get_length", "ft_array_set_length",
             "ft_array_mux", "ft_array_contains", "ft_add", "ft_reduce_isum", "ft_sub", "ft_scale", "ft_neg",
             "ft_mul", "ft_index", "ft_equal", "ft_not_equal", "ft_array_prescan", "ft_array_scan",
             "ft_array_reduce"]
    assert len(ref38) == 38
    for n in ref38:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ft.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_error_strings():
    assert ft.error_str(0) == "No error"
    assert ft.error_str(ft.FT_ERR_SIZE_MISMATCH) == "Size mismatch"
    assert ft.error_str(99) == "Unknown error"


def splitmix_at(seed, ctr):
    M = 2**64 - 1
    z = (seed + (ctr + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_random_scalars_definition():
    got = ft.random_scalars(40, seed=0xF7C00001, stream=3)
    for i in range(40):
        wide = 0
        for j in range(8):
            wide |= splitmix_at(0xF7C00001, (3 << 40) + 8 * i + j) << (64 * j)
        assert int.from_bytes(bytes(got[i]), "little") == wide % edref.L


def test_scalar_reduce64_matches_python_and_oracle():
    rng = np.random.default_rng(1)
    wide = rng.integers(0, 256, size=(500, 64), dtype=np.uint8)
    wide[0] = 0
    wide[1] = 255
    wide[2] = np.frombuffer(int.to_bytes(edref.L, 64, "little"), dtype=np.uint8)
    wide[3] = np.frombuffer(int.to_bytes(edref.L - 1, 64, "little"), dtype=np.uint8)
    got = ft.scalar_reduce64(wide)
    assert np.array_equal(got, O.scalar_reduce64(wide))
    for i in range(500):
        assert int.from_bytes(bytes(got[i]), "little") == int.from_bytes(bytes(wide[i]), "little") % edref.L
