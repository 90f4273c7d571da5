"""Textbook Ed25519 arithmetic in Python big integers (RFC 8032 section 5.1).

Used by the CPU tests to pin the oracle with closed forms: field operations
mod p, affine twisted-Edwards addition with a = -1, double-and-add scalar
multiplication, point encoding/decoding.  Pure Python, for small cases only.
"""
import hashlib

P = 2**255 - 19
L = 2**252 + 27742317777372353535851937790883648493
D = (-121665 * pow(121666, P - 2, P)) % P
SQRTM1 = pow(2, (P - 1) // 4, P)


def inv(x):
    return pow(x, P - 2, P)


def recover_x(y, sign):
    """x with x^2 = (y^2-1)/(dy^2+1); None when no root exists."""
    u = (y * y - 1) % P
    v = (D * y * y + 1) % P
    x2 = u * inv(v) % P
    if x2 == 0:
        return 0  # this reference decoder accepts x = 0 with either sign bit
    x = pow(x2, (P + 3) // 8, P)
    if (x * x - x2) % P != 0:
        x = x * SQRTM1 % P
    if (x * x - x2) % P != 0:
        return None
    if (x & 1) != sign:
        x = P - x
    return x


GY = 4 * inv(5) % P
GX = recover_x(GY, 0)
G = (GX, GY)
IDENT = (0, 1)


def add(p1, p2):
    x1, y1 = p1
    x2, y2 = p2
    t = D * x1 * x2 * y1 * y2 % P
    x3 = (x1 * y2 + y1 * x2) * inv(1 + t) % P
    y3 = (y1 * y2 + x1 * x2) * inv(1 - t) % P
    return (x3, y3)


def neg(p1):
    return ((-p1[0]) % P, p1[1])


def mul(k, pt):
    r = IDENT
    q = pt
    while k > 0:
        if k & 1:
            r = add(r, q)
        q = add(q, q)
        k >>= 1
    return r


def encode(pt):
    x, y = pt
    return int.to_bytes(y | ((x & 1) << 255), 32, "little")


def decode(s):
    """Reference decoding semantics: bit 255 is the sign, y taken mod p
    (values >= p accepted), failure when no square root exists."""
    v = int.from_bytes(s, "little")
    sign = v >> 255
    y = (v & ((1 << 255) - 1)) % P
    x = recover_x(y, sign)
    if x is None:
        return None
    return (x, y)


def fe_bytes(x):
    return int.to_bytes(x % P, 32, "little")


def scalar_bytes(k):
    return int.to_bytes(k, 32, "little")


def rfc8032_public_key(secret):
    h = hashlib.sha512(secret).digest()
    a = bytearray(h[:32])
    a[0] &= 248
    a[31] &= 127
    a[31] |= 64
    return encode(mul(int.from_bytes(bytes(a), "little"), G)), bytes(a)
