"""Generates mt_jump.inc: jump polynomials that move an MT19937 state ahead by whole chunks of its
output stream, so that host threads can each produce one chunk of a long stream.

MT19937 (Matsumoto and Nishimura, 1998) is linear over GF(2): its state, the window of the last 624
untempered words x_s .. x_{s+623}, advances one word at a time by a linear map A.  The minimal
polynomial phi of the word sequence has degree 19937 (the period exponent); it is found here by
the Berlekamp-Massey algorithm on 40,000 bits of one bit position of the sequence.  For a jump of
e words, P = x^e mod phi gives A^e W = P(A) W for every window W, up to the 31 low bits of the
window's oldest word, which no later word depends on.  Applying P(A) is Horner's rule: 19937
single-word steps of an accumulator window, adding W wherever P has a term.

The chunk length is L = 624 * 2^14 words; the file holds P_k = x^(k L) mod phi for k = 1 .. 15, each
as 312 64-bit words (bit i of the polynomial in word i / 64, bit i % 64).  Before writing, the
script checks a jump against direct generation.  Run from this directory:

    python3 gen_mt_jump.py > mt_jump.inc
"""
import sys

N, M = 624, 397
CHUNK_LOG2 = 14
CHUNK = N << CHUNK_LOG2
NPOLY = 15
WORDS = (19937 + 63) // 64


def mt_words(seed: int, count: int) -> list:
    """Untempered words x_0 .. x_{count-1} of mt19937(seed) (x_0 .. x_623 is the seed window)."""
    x = [seed & 0xFFFFFFFF]
    for i in range(1, N):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + i) & 0xFFFFFFFF)
    while len(x) < count:
        k = len(x) - N
        y = (x[k] & 0x80000000) | (x[k + 1] & 0x7FFFFFFF)
        x.append(x[k + M] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0))
    return x


def berlekamp_massey(bits) -> tuple:
    """Connection polynomial C (int, bit i = c_i, c_0 = 1) and length L of the shortest recurrence
    s_n = sum_{i=1..L} c_i s_{n-i} over GF(2)."""
    c, b, length, m, r = 1, 1, 0, 1, 0
    for n, s in enumerate(bits):
        r = (r << 1) | s  # bit i of r is s_{n-i}
        if (c & r).bit_count() & 1:
            t = c
            c ^= b << m
            if 2 * length <= n:
                length, b, m = n + 1 - length, t, 1
            else:
                m += 1
        else:
            m += 1
    return c, length


def polymod(p: int, phi: int, deg: int) -> int:
    while p.bit_length() > deg:
        p ^= phi << (p.bit_length() - 1 - deg)
    return p


def clmul(a: int, b: int) -> int:
    r = 0
    while b:
        low = b & -b
        r ^= a << (low.bit_length() - 1)
        b ^= low
    return r


def x_pow_mod(e_small: int, squarings: int, phi: int, deg: int) -> int:
    """x^(e_small * 2^squarings) mod phi."""
    p = polymod(1 << e_small, phi, deg)
    for _ in range(squarings):
        p = polymod(clmul(p, p), phi, deg)
    return p


def jump_window(w0: list, p: int) -> list:
    """P(A) W0 by Horner's rule on a linear buffer of words."""
    deg = p.bit_length() - 1
    a = [0] * (N + deg + 1)
    base = 0
    for i in range(deg, -1, -1):
        if i != deg:  # acc = A acc (the first step acts on zero)
            y = (a[base] & 0x80000000) | (a[base + 1] & 0x7FFFFFFF)
            a[base + N] = a[base + M] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            base += 1
        if (p >> i) & 1:
            for k in range(N):
                a[base + k] ^= w0[k]
    return a[base:base + N]


def main() -> None:
    seq = mt_words(5489, N + 40000)
    bits = [(w & 1) for w in seq[N:N + 40000]]
    c, length = berlekamp_massey(bits)
    if length != 19937:
        raise SystemExit(f"unexpected recurrence length {length}")
    phi = 0
    for i in range(length + 1):  # phi(x) = x^L C(1/x)
        if (c >> i) & 1:
            phi |= 1 << (length - i)
    # check: a jump of 3 blocks plus 5 words against direct generation
    e = 3 * N + 5
    w = jump_window(seq[:N], polymod(1 << e, phi, length))
    direct = seq[e:e + N]
    if w[1:] != direct[1:] or (w[0] ^ direct[0]) & 0x80000000:
        raise SystemExit("jump check failed")
    polys = [x_pow_mod(N * k, CHUNK_LOG2, phi, length) for k in range(1, NPOLY + 1)]
    out = sys.stdout
    out.write("// Generated by gen_mt_jump.py: MT19937 jump polynomials x^(k L) mod phi, k = 1..15, for\n")
    out.write(f"// chunks of L = {CHUNK} words; {WORDS} little-endian 64-bit words each.\n")
    out.write(f"constexpr int64_t kMtJumpChunk = {CHUNK};\n")
    out.write(f"constexpr int kMtJumpPolys = {NPOLY};\n")
    out.write(f"constexpr int kMtJumpWords = {WORDS};\n")
    out.write("alignas(64) static const uint64_t kMtJumpPoly[kMtJumpPolys][kMtJumpWords] = {\n")
    for p in polys:
        ws = [(p >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(WORDS)]
        out.write("  {")
        for i, v in enumerate(ws):
            if i % 6 == 0:
                out.write("\n   ")
            out.write(f" 0x{v:016x}ull,")
        out.write("\n  },\n")
    out.write("};\n")


if __name__ == "__main__":
    main()
