"""padded_size: the gridded size of a padded image (WSClean's -padding), a
multiple of the uv tile with only the radices 2, 3, 5 and 7 (CPU only)."""
from wsclean_amd.accuracy import padded_size


def _smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_padded_sizes():
    assert padded_size(1024, 1.0) == 1024
    assert padded_size(1024, 1.25) == 1280
    assert padded_size(1024, 1.5) == 1536
    assert padded_size(4096, 1.2) == 5040   # 4928 and the multiples of 16 between hold 11 or 13
    assert padded_size(4096, 1.5) == 6144
    for w in (512, 1000, 1024, 2048, 4096):
        for p in (1.0, 1.1, 1.2, 1.3, 1.5, 2.0):
            n = padded_size(w, p)
            assert n >= w * p - 1e-6 and n % 16 == 0 and _smooth(n)
            assert all(not _smooth(m) for m in range(n - 16, int(w * p) - 1, -16)
                       if m >= w * p)
