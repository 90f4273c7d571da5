"""The schedule table's host API (no device work): set / get, size classes,
save / load round trip, and rejection of out-of-range or malformed input
(including a file in an older, shorter format)."""
import pytest

import libftcrypto_amd as ft


@pytest.fixture(autouse=True)
def _keep():
    saved = ft.tune_table()
    yield
    for c, v in saved.items():
        ft.tune_set(c, *v)


def test_size_classes():
    assert [ft.size_class(n) for n in (0, 1, 2**14 - 1, 2**14, 2**18 - 1, 2**18, 2**22 - 1, 2**22, 2**30)] == \
        [0, 0, 0, 1, 1, 2, 2, 3, 3]


def test_set_get_save_load(tmp_path):
    ft.tune_set(0, 5, 4, 4, 512, 3, 4)
    ft.tune_set(3, 6, 2, 5, 4096, 4, 3)
    t = ft.tune_table()
    assert t[0] == (5, 4, 4, 512, 3, 4) and t[3] == (6, 2, 5, 4096, 4, 3)
    p = str(tmp_path / "t.txt")
    ft.tune_save(p)
    ft.tune_set(0, 4, 4, 4, 2048, 4, 4)
    ft.tune_load(p)
    assert ft.tune_table() == t


@pytest.mark.parametrize("bad", [(0, 3, 4, 4, 2048, 4, 4), (0, 8, 4, 4, 2048, 4, 4), (0, 4, 4, 8, 2048, 4, 4),
                                 (0, 4, 5, 4, 2048, 4, 4), (0, 4, 0, 4, 2048, 4, 4),
                                 (0, 4, 4, 4, 0, 4, 4),
                                 (0, 4, 4, 4, 4097, 4, 4), (0, 4, 4, 4, 2048, 5, 4), (0, 4, 4, 4, 2048, 2, 4),
                                 (0, 4, 4, 4, 2048, 4, 5)])
def test_set_rejects_out_of_range(bad):
    before = ft.tune_table()
    with pytest.raises(ft.FtError) as e:
        ft.tune_set(*bad)
    assert e.value.code == ft.FT_ERR_BAD_SIZE
    assert ft.tune_table() == before
    with pytest.raises(ft.FtError) as e:
        ft.tune_set(4, 4, 4, 4, 2048, 4, 4)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS


@pytest.mark.parametrize("text", ["0 4 4 4 2048 4 4\n1 4 4 4 2048 4 4\n2 4 4 4 2048 4 4\n",
                                  "0 4 4 4 2048 4 4\n1 4 4 4 2048 4 4\n2 4 4 4 2048 4 4\n3 4 9 4 2048 4 4\n",
                                  "0 4 4 4 2048\n1 4 4 4 2048\n2 4 4 4 2048\n3 4 4 4 2048\n",
                                  "garbage\n"])
def test_load_rejects_malformed(tmp_path, text):
    before = ft.tune_table()
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(ft.FtError) as e:
        ft.tune_load(str(p))
    assert e.value.code == ft.FT_ERR_BAD_SIZE
    assert ft.tune_table() == before


def test_ct_lanes_set_get_save_load(tmp_path):
    before = ft.ct_lanes_table()
    try:
        ft.tune_set_ct_lanes(0, 1)
        ft.tune_set_ct_lanes(3, 2)
        assert ft.ct_lanes_table()[0] == 1 and ft.ct_lanes_table()[3] == 2
        p = str(tmp_path / "t.txt")
        ft.tune_save(p)
        assert open(p).read().splitlines()[1].split()[7] == "1"
        ft.tune_set_ct_lanes(0, 2)
        ft.tune_load(p)
        assert ft.ct_lanes_table()[0] == 1 and ft.ct_lanes_table()[3] == 2
        for bad in (0, 3):
            with pytest.raises(ft.FtError) as e:
                ft.tune_set_ct_lanes(1, bad)
            assert e.value.code == ft.FT_ERR_BAD_SIZE
        # a file without the lanes column (the older format) loads with the
        # defaults: four waves per 64 ciphertexts below 2^18, two above
        q = tmp_path / "old.txt"
        q.write_text("".join("%d 4 1 4 512 3 4\n" % c for c in range(4)))
        ft.tune_load(str(q))
        assert ft.ct_lanes_table() == {0: 4, 1: 4, 2: 2, 3: 2}
    finally:
        for c, v in before.items():
            ft.tune_set_ct_lanes(c, v)


def test_scan_lookback_set_get_save_load(tmp_path):
    before = ft.scan_lookback_table()
    try:
        ft.tune_set_scan_lookback(2, 1)
        assert ft.scan_lookback_table()[2] == 1
        p = str(tmp_path / "t.txt")
        ft.tune_save(p)
        rows = [ln.split() for ln in open(p).read().splitlines()[1:]]
        assert all(len(r) == 9 for r in rows) and rows[2][8] == "1"
        ft.tune_set_scan_lookback(2, 0)
        ft.tune_load(p)
        assert ft.scan_lookback_table()[2] == 1
        with pytest.raises(ft.FtError) as e:
            ft.tune_set_scan_lookback(1, 2)
        assert e.value.code == ft.FT_ERR_BAD_SIZE
        # files of eight columns (written before the scan form existed) load with the defaults
        q = tmp_path / "eight.txt"
        q.write_text("".join("%d 4 1 4 512 3 4 2\n" % c for c in range(4)))
        ft.tune_load(str(q))
        assert ft.scan_lookback_table() == {0: 0, 1: 0, 2: 0, 3: 0}
    finally:
        for c, v in before.items():
            ft.tune_set_scan_lookback(c, v)
