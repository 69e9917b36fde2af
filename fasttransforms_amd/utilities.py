"""Test and benchmark utilities of the reference (src/ftutilities.c), in numpy.

* ``fill``: i.i.d. U(-1, 1) entries in each layout's valid triangle, zeros
  elsewhere -- sphrand / trirand / diskrand (ftutilities.c:75-113) with a
  seeded, counter-based SplitMix64 stream instead of libc rand(), so inputs
  are reproducible across hosts and languages: entry k = i + n*j of stream
  `seed` is the k-th SplitMix64 output, u = (top 53 bits) * 2^-53, value 2u-1.
* ``ones``: sphones / triones / diskones (ftutilities.c:67-105).
* ``rel2``, ``relinf``: ft_norm_2arg / ft_norm_1arg and the inf-norm variants
  (ftutilities_source.c:6-36) as relative errors.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def rand_entries(seed, k):
    """U(-1,1) entries k (uint64 array) of stream `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = _mix64(z)
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0


# the layout of each executor kind's matrices: sphv is the sphere's, the
# rectangle-disk (full width M = n) keeps rows i < n - j as the triangle does
# (rectdiskrand, ftutilities.c:115-129), the annulus is the disk's
_LAYOUT = {"sph": "sph", "sphv": "sph", "tri": "tri", "rectdisk": "tri", "disk": "disk", "annulus": "disk"}


def valid_rows(kind, n, j):
    j = np.asarray(j)
    kind = _LAYOUT[kind]
    if kind == "sph":
        return n - (j + 1) // 2
    if kind == "tri":
        return n - j
    return n - (j + 3) // 4


def full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[_LAYOUT[kind]]


def _mask(kind, n, M):
    rows = valid_rows(kind, n, np.arange(M))
    return np.arange(n)[:, None] < rows[None, :]


def fill(kind, n, M, seed):
    """Seeded valid-triangle fill; Fortran-ordered (n, M) float64."""
    k = np.arange(n * M, dtype=np.uint64).reshape(M, n).T  # k = i + n*j
    A = np.where(_mask(kind, n, M), rand_entries(seed, k), 0.0)
    return np.asfortranarray(A)


def ones(kind, n, M):
    return np.asfortranarray(_mask(kind, n, M).astype(np.float64))


def rel2(A, B):
    d = np.linalg.norm((np.asarray(A, dtype=np.float64) - np.asarray(B, dtype=np.float64)).ravel())
    b = np.linalg.norm(np.asarray(B, dtype=np.float64).ravel())
    return d / b if b else d


def relinf(A, B):
    d = np.max(np.abs(np.asarray(A, dtype=np.float64) - np.asarray(B, dtype=np.float64)))
    b = np.max(np.abs(np.asarray(B, dtype=np.float64)))
    return d / b if b else d


def rotations(kind, n, M=None):
    """Number of rotations ft_execute_<kind>_hi2lo applies to an n x M matrix
    (6 algorithmic flops each).  Column col holds order m: tri and rectdisk
    m = col, sph and disk (annulus) m = (col+1)//2, sphv m = (col-1)//2 up to
    M//2 - 1; orders below 2 (tri: 1) apply nothing."""
    if M is None:
        M = full_M(kind, n)
    geo = {"sphv": "sph", "rectdisk": "sph", "annulus": "disk"}.get(kind, kind)
    step = 1 if geo == "tri" else 2
    total = 0
    for col in range(1, M):
        if kind == "tri":
            m, m1 = col, 0
        else:
            m = col if kind == "rectdisk" else (col - 1) // 2 if kind == "sphv" else (col + 1) // 2
            if m < 2 or (kind == "sphv" and m > M // 2 - 1):
                continue
            m1 = m % 2
        jmax = m - step
        for j in range(m1, jmax + 1, step):
            top = {"sph": n - 3 - j, "tri": n - 2 - j, "disk": n - 2 - (j + 1) // 2}[geo]
            if top >= 0:
                total += top + 1
    return total


def tet_rotations(N, L, M):
    """Number of rotations ft_execute_tet_hi2lo applies to an N x L x M array
    (drivers.c:510-533; 6 algorithmic flops each): in slab m, column l < L-m
    takes the triangle's orders l+m-1 .. 0 (N-1-j rotations each), and the
    block rotations take, for j < m with j <= L-2, L-1-j rotations of whole
    rows (N entries each)."""
    total = 0
    for m in range(M):
        for l in range(max(L - m, 0)):
            e = min(l + m, N - 1)
            total += e * (N - 1) - e * (e - 1) // 2
        jj = min(m, L - 1) if L >= 1 else 0
        total += N * (jj * (L - 1) - jj * (jj - 1) // 2)
    return total
