"""CPU pins of the oracle's restatement of Thermalizing_MC_DEV (MD_DiffScheme_GPU.F90:1451-1805)
fed by Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).

- the generator: the known-answer blocks of Philox4x32-10 published with the Random123 library
  (zero counter and key, all-ones counter and key, and the pi-digit counter/key pair);
- the uniforms lie in (0, 1] and are uniform (Kolmogorov-Smirnov);
- the transformation: each free component is Gaussian with variance kB T / m (KS test on the
  normalised components), every box has zero total momentum afterwards, components with a
  FIXVEL or FIXPOS bit keep their value apart from the box's momentum shift, inactive atoms end
  at minus the box's mean velocity, draws differ, and the same (seed, draw) reproduces exactly."""
import numpy as np
import pytest

stats = pytest.importorskip("scipy.stats")

KB = 8.617333262e-5


def test_philox_known_answers(orc):
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        assert tuple(orc.philox4x32_10(ctr, key)) == want


def _case(n=20000, nbox=2, seed=7):
    rng = np.random.default_rng(seed)
    natoms = [n // 2, n - n // 2][:nbox]
    ityp = rng.integers(1, 3, size=n).astype(np.int32)
    cm = np.array([183.84, 63.5]) * 103.6427
    statu = np.ones(n, dtype=np.int32)
    return natoms, ityp, statu, cm


def test_maxwell_components_and_zero_momentum(orc):
    natoms, ityp, statu, cm = _case()
    temp = 700.0
    v = orc.thermalize_philox(natoms, ityp, statu, cm, temp * KB, 0x1234ABCD5678, 0,
                              np.zeros((len(ityp), 3)))
    m = cm[ityp - 1]
    # each component, scaled by sqrt(m / kT), is standard normal (momentum removal shifts it by
    # O(1/sqrt(N)), far below what the test resolves)
    z = (v * np.sqrt(m / (KB * temp))[:, None]).ravel()
    assert stats.kstest(z, "norm").pvalue > 1e-3
    off = 0
    for nb in natoms:
        s = slice(off, off + nb)
        p = (m[s, None] * v[s]).sum(axis=0)
        assert np.max(np.abs(p)) <= 1e-12 * (m[s, None] * np.abs(v[s])).sum()
        off += nb
    # temperature within sampling error (3N degrees of freedom less the momentum)
    t = (m[:, None] * v ** 2).sum() / (3 * len(ityp) * KB)
    assert abs(t / temp - 1.0) < 5 * np.sqrt(2.0 / (3 * len(ityp)))


def test_uniforms_are_open_at_zero_and_uniform(orc):
    u = []
    for i in range(4000):
        r = orc.philox4x32_10((i, 1, 0, 0), (99, 5))
        for lo, hi in ((r[0], r[1]), (r[2], r[3])):
            u.append((((hi & 0x1FFFFF) << 32 | lo) + 1) * 2.0 ** -53)
    u = np.array(u)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def test_fixed_components_inactive_atoms_and_draws(orc):
    """Kept components and an inactive atom's zero velocity both move by minus the box's mean
    velocity VT: v[FIX] = 0.25 - VT and v[inactive] = -VT, component by component."""
    natoms, ityp, statu, cm = _case(n=3000, nbox=2)
    statu = statu.copy()
    statu[10] = 1 | 16          # FIXVELX
    statu[11] = 1 | 4 | 64      # FIXPOSY, FIXVELZ
    statu[12] = 0               # inactive
    v0 = np.full((len(ityp), 3), 0.25)
    v = orc.thermalize_philox(natoms, ityp, statu, cm, 300.0 * KB, 42, 3, v0)
    vt = -v[12]
    assert np.all(vt != 0.0)
    assert v[10, 0] == 0.25 - vt[0]
    assert v[11, 1] == 0.25 - vt[1] and v[11, 2] == 0.25 - vt[2]
    assert v[10, 1] != 0.25 - vt[1]  # a free component was drawn
    m = cm[ityp - 1]
    s = slice(0, natoms[0])
    assert np.max(np.abs((m[s, None] * v[s]).sum(axis=0))) <= 1e-12 * (m[s, None] * np.abs(v[s])).sum()
    # same (seed, draw) reproduces, another draw differs
    v2 = orc.thermalize_philox(natoms, ityp, statu, cm, 300.0 * KB, 42, 3, v0)
    assert np.array_equal(v, v2)
    v3 = orc.thermalize_philox(natoms, ityp, statu, cm, 300.0 * KB, 42, 4, v0)
    assert np.mean(np.abs(v[20:] - v3[20:])) > 0.1 * np.mean(np.abs(v[20:]))
