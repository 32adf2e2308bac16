"""The one-wave pair sweep's lane schedules (surfmol_amd/csrc/sweep_schedules.h),
checked on the CPU: every unordered pair of atoms is summed exactly once, and
a replay of the sweep's register rotation (lane v holds on round t the
accumulator that started in lane v + t - 1, as DPP wave_rol:1 moves it) with
the schedule's source lists gives each atom the sum of its pair forces, for
every size the schedules cover and with 1-2 exclusions dropped."""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "surfmol_amd", "csrc", "sweep_schedules.h")


def load():
    txt = open(HDR).read()
    nmax = int(re.search(r"kSchedMaxAtoms = (\d+);", txt).group(1))
    tmax = [int(x) for x in re.search(r"kSchedTmax\[\d+\] = \{([^}]*)\}", txt).group(1).split(",")]
    body = txt[txt.index("kSched[") :]
    ents = re.findall(r"\{(\d+),\s*\{([^}]*)\},\s*\{([^}]*)\},\s*\{([^}]*)\},\s*\{([^}]*)\}\}", body)
    assert len(ents) == nmax
    out = {}
    for n, (T, a, g, keep, src) in enumerate(ents, start=1):
        out[n] = dict(T=int(T), a=[int(x) for x in a.split(",")], g=[int(x) for x in g.split(",")],
                      keep=[int(x.rstrip("u"), 16) for x in keep.split(",")],
                      src=[int(x.rstrip("u"), 16) for x in src.split(",")])
    return nmax, tmax, out


NMAX, TMAX, SCHED = load()


def test_schedule_shapes():
    for n, s in SCHED.items():
        assert len(s["a"]) == 64 and len(s["g"]) == 64
        assert s["a"][:n] == list(range(n)) and s["g"][:n] == list(range(n))
        assert all(-1 <= x < n for x in s["a"] + s["g"])
        assert 0 <= s["T"] <= 32 and s["T"] <= max(1, n // 2)
        assert TMAX[n] == max(SCHED[m]["T"] for m in range(1, n + 1))


@pytest.mark.parametrize("n", list(range(1, NMAX + 1)))
def test_every_pair_summed_once(n):
    s = SCHED[n]
    T, a, g, keep = s["T"], s["a"], s["g"], s["keep"]
    seen = {}
    for v in range(64):
        assert keep[v] >> T == 0
        for t in range(1, T + 1):
            if not (keep[v] >> (t - 1)) & 1:
                continue
            i, j = a[v], g[(v + t) & 63]
            assert i >= 0 and j >= 0 and i != j
            key = (min(i, j), max(i, j))
            assert key not in seen
            seen[key] = (v, t)
    assert len(seen) == n * (n - 1) // 2


def replay(n, bonds, rng):
    """Per-atom pair forces from the schedule, as the kernel sums them."""
    s = SCHED[n]
    T, a, g, keep, src = s["T"], s["a"], s["g"], s["keep"], s["src"]
    F = rng.normal(size=(n, n, 3))
    F = F - F.transpose(1, 0, 2)                       # F[i, j] = -F[j, i]
    bonded = np.zeros((n, n), bool)
    for i, j in bonds:
        bonded[i, j] = bonded[j, i] = True
    own = np.zeros((64, 3))
    acc = np.zeros((64, 3))
    for t in range(1, T + 1):
        if t > 1:
            acc = np.roll(acc, -1, axis=0)             # lane v takes lane v+1's value
        for v in range(64):
            i, j = a[v], g[(v + t) & 63]
            if not (keep[v] >> (t - 1)) & 1 or bonded[i, j]:
                continue
            own[v] += F[i, j]
            acc[v] -= F[i, j]
    got = np.zeros((n, 3))
    for i in range(n):
        tot = own[i].copy()
        for k in range(4):
            b = (src[i] >> (8 * k)) & 0xFF
            if b == 0xFF:
                break
            tot += acc[b >> 1] if b & 1 else own[b >> 1]
        got[i] = tot
    want = np.where(bonded[:, :, None], 0.0, F).sum(axis=1)
    return got, want


@pytest.mark.parametrize("n", list(range(2, NMAX + 1)))
def test_replay_gives_pair_sums(n):
    rng = np.random.default_rng(1000 + n)
    # a chain plus a few rings' worth of random bonds (1-2 exclusions)
    bonds = [(i, i + 1) for i in range(n - 1)]
    bonds += [tuple(sorted(rng.choice(n, 2, replace=False))) for _ in range(n // 4)]
    got, want = replay(n, bonds, rng)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)


def test_ptcda_takes_13_rounds():
    # the uniform batch's 38-atom molecule: 13 rounds where rows 0..N/2
    # repeated after the atoms would take 19
    assert SCHED[38]["T"] <= 13
