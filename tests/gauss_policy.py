"""The comparison rule for Gaussian draws of the device against the reference.

The device's Box-Muller uses correctly rounded log and cos; the reference uses
host glibc's, which misround on about 1 draw in 600. A device value must
therefore equal the reference value, and where it does not, the reference
value must be glibc's draw and the device value the correctly rounded one
(log and cos in quad precision, rounded to double). Anything else fails.

Key layouts (the reference's, fields.hpp:49-66, scalar_field.hpp:33-51):
link momenta P[(s*D + mu)*ngen + a] use key(stream, s, mu, a); scalar
components phi[2*(s*d + k) + part] use key(stream, s, 2k + part)."""
import numpy as np

import oracle_lib as O

TAG_P, TAG_PI, TAG_P_PURE = 0xC3, 0xD4, 0xA1


def momentum_keys(idx, D, ngen, stream):
    s, rem = np.divmod(idx, D * ngen)
    mu, a = np.divmod(rem, ngen)
    return [O.key(stream, int(x), int(m), int(b)) for x, m, b in zip(s, mu, a)]


def scalar_keys(idx, d, stream):
    c, part = np.divmod(idx, 2)
    s, k = np.divmod(c, d)
    return [O.key(stream, int(x), int(2 * kk + p)) for x, kk, p in zip(s, k, part)]


def assert_equal_but_glibc_misrounds(dev, ref, keys_of, seed, scale=1.0, what=""):
    """dev == ref everywhere, except at draws where glibc misrounds: there ref is
    scale * (glibc draw) and dev is scale * (correctly rounded draw). Returns the
    number of such exceptions."""
    dev = np.ascontiguousarray(dev, dtype=np.float64).ravel()
    ref = np.ascontiguousarray(ref, dtype=np.float64).ravel()
    assert dev.shape == ref.shape
    idx = np.nonzero(dev != ref)[0]
    if idx.size == 0:
        return 0
    assert idx.size <= max(8, dev.size // 100), f"{what}: {idx.size} of {dev.size} draws differ"
    keys = keys_of(idx)
    q = scale * O.gauss_quad_many(seed, keys)
    gl = scale * O.gauss_many(seed, keys)
    assert np.array_equal(ref[idx], gl), f"{what}: the reference value is not glibc's draw"
    assert np.array_equal(dev[idx], q), f"{what}: device value differs from the correctly rounded draw at {idx[:8]}"
    assert np.all(q != gl)
    return int(idx.size)
