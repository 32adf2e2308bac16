"""Hand-derived pins of the oracle's integrator steps (CPU).

FIRE follows the loop of surfmol/tests/relax_pentacene_uff.rs:142-193 with
FireState semantics (molff/src/raff.rs:1040-1062) and the |F| clamp of
step_fire (raff.rs:1089-1096); damped MD follows MolWorld::run_md
(surfmol/src/mol_world.rs:141-164) and DynamicAtoms::move_atom_md
(moltopo/src/molecular.rs:118-137).

The expected values below were derived by hand, in exact rational
arithmetic, and are written out as numbers.  They are chosen so that each
plausible mistake in the step order fails:
  * taking v.F over the velocity before the update (the first step would
    then be "uphill": v = 0 gives v.F = 0),
  * moving with the dt of the next step after the n_min transition,
  * summing the clamped |F|^2 in the mixing norm,
  * purging the velocities before the move in run_md, or summing v.F over
    the updated velocity.
"""
import numpy as np
import pytest

# One atom in the 1-D harmonic well F = -x (along x), x0 = 1, v0 = 0, with
# dt0 = 0.1, dt_max = 0.15 and the default FireState constants
# (alpha0 = 0.1, n_min = 5, f_inc = 1.1, f_dec = 0.5, f_alpha = 0.99).
# In one dimension a downhill mixing (1-a) v + a |v| F/|F| returns v itself,
# so the recursion per step is
#   v += F dt;  downhill (v F > 0): n_pos += 1 and, once n_pos > 5,
#   dt = min(1.1 dt, 0.15), a *= 0.99;  uphill: n_pos = 0, dt /= 2, a = 0.1,
#   v = 0;  x += v dt_old.
# Rows: (step, x, v, dt, alpha, n_pos) after the step.
FIRE_1D = [
    (1, 0.99, -0.1, 0.1, 0.1, 1),
    (2, 0.9701, -0.199, 0.1, 0.1, 2),
    (3, 0.940499, -0.29601, 0.1, 0.1, 3),
    (4, 0.90149301, -0.3900599, 0.1, 0.1, 4),
    (5, 0.8534720899, -0.480209201, 0.1, 0.1, 5),
    # n_pos = 6 > n_min: dt grows for the NEXT step, this one moves with 0.1
    (6, 0.796916448901, -0.56555640999, 0.11, 0.099, 6),
    (7, 0.7250625547703979, -0.65321721936911, 0.121, 0.09801, 7),
    (8, 0.6354076303623422, -0.7409497884963282, 0.1331, 0.0970299, 8),
    (9, 0.5255305797429575, -0.8255225440975559, 0.14641, 0.096059601, 9),
    (10, 0.3934006093611362, -0.9024654762777223, 0.15, 0.09509900499, 10),   # capped at dt_max
    (11, 0.24917927420885227, -0.9614755676818927, 0.15, 0.0941480149401, 11),
    (12, 0.0993514053868692, -0.9988524588132206, 0.15, 0.093206534790699, 12),
    (13, -0.05271187005631845, -1.013755169621251, 0.15, 0.09227446944279201, 13),
    # past the minimum: v.F < 0 -> reset, x does not move
    (14, -0.05271187005631845, 0.0, 0.075, 0.1, 0),
    (15, -0.052415365787251665, 0.003953390254223884, 0.075, 0.1, 1),
    (16, -0.05182402508563158, 0.007884542688267758, 0.075, 0.1, 2),
    (17, -0.05094117424290482, 0.011771344569690128, 0.075, 0.1, 3),
    (18, -0.04977177929506172, 0.015591932637907988, 0.075, 0.1, 4),
    (19, -0.0483224180886839, 0.019324816085037617, 0.075, 0.1, 5),
    (20, -0.04660124328055723, 0.02294899744168891, 0.0825, 0.099, 6),
    (21, -0.0443907712795396, 0.02679360001233488, 0.09075, 0.09801, 7),
    (22, -0.04159366910712436, 0.0308220625059531, 0.099825, 0.0970299, 8),
    (23, -0.03810237453100898, 0.03497415052457179, 0.1098075, 0.096059601, 9),
    (24, -0.033802523989087126, 0.039158077015885556, 0.12078825, 0.09509900499, 10),
    (25, -0.028579516283247217, 0.043241024734110414, 0.132867075, 0.0941480149401, 11),
    (26, -0.02232967475428161, 0.04703830146758034, 0.1461537825, 0.093206534790699, 12),
    (27, -0.014977866494624709, 0.05030186789491336, 0.15, 0.09227446944279201, 13),
    (28, -0.00709558431425865, 0.052548547869107065, 0.15, 0.09135172474836409, 14),
    (29, 0.0009463485131782296, 0.05361288551624586, 0.15, 0.09043820750088045, 15),
    (30, 0.0009463485131782296, 0.0, 0.075, 0.1, 0),
]


def test_fire_1d_well_sequence(oracle):
    fp = oracle.FireParams(dt0=0.1, dt_max=0.15)
    x = np.array([[1.0, 0.0, 0.0]])
    v = np.zeros((1, 3))
    state = np.array([fp.dt0, fp.alpha0, 0.0])
    for (step, xe, ve, dte, ae, npe) in FIRE_1D:
        f = -x * np.array([1.0, 0.0, 0.0])
        x, v, _, f2max = oracle.fire_step(f, x, v, state, fp)
        assert f2max == pytest.approx(float(f[0, 0]) ** 2, rel=1e-15)
        assert x[0, 0] == pytest.approx(xe, abs=1e-13), step
        assert v[0, 0] == pytest.approx(ve, abs=1e-13), step
        assert state[0] == pytest.approx(dte, rel=1e-13), step
        assert state[1] == pytest.approx(ae, rel=1e-13), step
        assert state[2] == npe, step
        assert x[0, 1] == 0.0 and x[0, 2] == 0.0 and v[0, 1] == 0.0


def test_fire_mixing_and_clamp_one_step(oracle):
    """Velocity mixing off the force direction, without and with the clamp.

    dt = 0.5, v_old = (3, 3, 0), x = 0, alpha = 0.1.
    * No clamp, F = (0, 2, 0): v = (3, 4) -> |v| = 5, |F| = 2,
      v_mix = 0.9 (3, 4) + 0.1 * 5 * (0, 1) = (2.7, 4.1); x = 0.5 v_mix.
    * flim = 2, F = (0, 6, 0): the clamped (0, 2) drives v = (3, 4) and
      v.F = 8 > 0, but |F| in the mixing is the unclamped 6 (step_fire sums
      f2 before clamping), so F_hat = (0, 2)/6:
      v_mix = (2.7, 3.6 + 0.5/3) = (2.7, 3.7666...); the returned force is
      the clamped one."""
    fp = oracle.FireParams(dt0=0.5, dt_max=1.0)
    v0 = np.array([[3.0, 3.0, 0.0]])
    x0 = np.zeros((1, 3))
    st = np.array([0.5, 0.1, 0.0])
    x, v, fc, f2 = oracle.fire_step(np.array([[0.0, 2.0, 0.0]]), x0, v0, st, fp)
    assert v[0] == pytest.approx([2.7, 4.1, 0.0], abs=1e-14)
    assert x[0] == pytest.approx([1.35, 2.05, 0.0], abs=1e-14)
    assert list(st) == [0.5, 0.1, 1.0]
    fpc = oracle.FireParams(dt0=0.5, dt_max=1.0, flim=2.0)
    st = np.array([0.5, 0.1, 0.0])
    x, v, fc, f2 = oracle.fire_step(np.array([[0.0, 6.0, 0.0]]), x0, v0, st, fpc)
    assert f2 == 36.0                                   # max |F|^2 of the unclamped force
    assert fc[0] == pytest.approx([0.0, 2.0, 0.0], abs=1e-15)
    assert v[0] == pytest.approx([2.7, 3.6 + 0.5 / 3.0, 0.0], abs=1e-14)
    assert x[0] == pytest.approx([1.35, 0.5 * (3.6 + 0.5 / 3.0), 0.0], abs=1e-14)


def test_fire_two_atoms_uphill_reset(oracle):
    """Two atoms: sums run over atoms.  v_old = (1,0,0) and (0,0,0), F =
    (-3,0,0) and (0,1,0), dt = 0.5: v = (-0.5,0,0), (0,0.5,0); v.F = 1.5 +
    0.5 = 2 > 0 -> downhill, |v| = sqrt(0.5), |F| = sqrt(10).  With F = (-5,0,0)
    on atom 0 instead: v = (-1.5,0,0), v.F = 7.5 + 0.5 > 0 as well; with v_old
    = (4,0,0): v = (1.5,0,0), v.F = -7.5 + 0.5 < 0 -> reset: v = 0, x unmoved,
    dt = 0.25, n_pos = 0, alpha = alpha0."""
    fp = oracle.FireParams(dt0=0.5, dt_max=1.0)
    x0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    st = np.array([0.5, 0.05, 3.0])
    f = np.array([[-3.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    x, v, _, _ = oracle.fire_step(f, x0, np.array([[1.0, 0, 0], [0, 0, 0]]), st, fp)
    vm, fm = np.sqrt(0.5), np.sqrt(10.0)
    ve = 0.95 * np.array([[-0.5, 0, 0], [0, 0.5, 0]]) + 0.05 * vm * f / fm
    assert v == pytest.approx(ve, abs=1e-15)
    assert x == pytest.approx(x0 + 0.5 * ve, abs=1e-15)
    assert list(st) == [0.5, 0.05, 4.0]
    st = np.array([0.5, 0.05, 3.0])
    f = np.array([[-5.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    x, v, _, _ = oracle.fire_step(f, x0, np.array([[4.0, 0, 0], [0, 0, 0]]), st, fp)
    assert (v == 0.0).all() and (x == x0).all()
    assert list(st) == [0.25, 0.1, 0.0]


def test_run_md_one_step_closed_form(oracle):
    """move_atom_md per atom, then the purge (mol_world.rs:145-160).

    dt = 0.1, cdamp = 0.5, flim = 2.5.
    atom 0: v = (-1,0,0), F = (3,4,0): |F| = 5 > flim -> F_c = (1.5,2,0);
            v' = 0.5 v + 0.1 F_c = (-0.35, 0.2, 0); x' = x + 0.1 v'.
    atom 1: v = (0,0,1), F = (0,0,1): v' = (0,0,0.6); x' = x + (0,0,0.06).
    Sums use the velocity BEFORE the update and the unclamped F:
      v.F = -3 + 1 = -2, |v|^2 = 2, |F|^2 = 25 + 1 = 26.
    v.F < 0, so every velocity is zeroed -- after both atoms moved.
    With v_0 = (+1,0,0) instead, v.F = 3 + 1 > 0 and the velocities stay."""
    x0 = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    f = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 1.0]])
    x, v, sums = oracle.md_step(f, x0, np.array([[-1.0, 0, 0], [0, 0, 1.0]]), 0.1, 2.5, 0.5)
    assert sums == pytest.approx((-2.0, 2.0, 26.0), abs=1e-15)
    assert x == pytest.approx(np.array([[-0.035, 0.02, 0.0], [2.0, 0.0, 0.06]]), abs=1e-15)
    assert (v == 0.0).all()
    x, v, sums = oracle.md_step(f, x0, np.array([[1.0, 0, 0], [0, 0, 1.0]]), 0.1, 2.5, 0.5)
    assert sums == pytest.approx((4.0, 2.0, 26.0), abs=1e-15)
    assert v == pytest.approx(np.array([[0.65, 0.2, 0.0], [0.0, 0.0, 0.6]]), abs=1e-15)
    assert x == pytest.approx(np.array([[0.065, 0.02, 0.0], [2.0, 0.0, 0.06]]), abs=1e-15)


def test_relax_fire_and_run_md_drive_the_steps(oracle, params):
    """orc_relax_fire / orc_run_md are the step functions above driven by the
    force field: the H2 bond (k = 100, l0 = 0.9) stretched to 1.0 A gives
    F = +-20 along x on the first evaluation (test_uff_cl.rs:8-56), so the
    first FIRE step from rest with dt0 = 0.01 gives v = (+-0.2, 0, 0) (1-D:
    mixing is the identity) and x = (0.002, 0.998); one run_md step with
    cdamp = 1 - 0.1 and flim = 50 gives the same move."""
    O = oracle
    s = O.OracleSystem(
        elems=["H", "H"], apos=np.array([[0.0, 0, 0], [1.0, 0, 0]]), bonds=np.array([[0, 1]], dtype=np.int32),
        angles=np.zeros((0, 3), np.int32), dihedrals=np.zeros((0, 4), np.int32), inversions=np.zeros((0, 4), np.int32),
        neighs=np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32),
        neigh_bs=np.array([[0, -1, -1, -1], [0, -1, -1, -1]], np.int32), types=["H_", "H_"],
        bon_params=np.array([[100.0, 0.9]]), ang_params=np.zeros((0, 5)), dih_params=np.zeros((0, 3)),
        inv_params=np.zeros((0, 4)), ang_ngs=np.zeros((0, 2), np.int32), dih_ngs=np.zeros((0, 3), np.int32),
        inv_ngs=np.zeros((0, 3), np.int32), reqs=np.zeros((2, 4)), nb_mode=O.NB_OFF)
    r = O.relax_fire(s, O.FireParams(dt0=0.01, dt_max=0.1), 1, 1e-3)
    assert r["steps"] == 1 and not r["converged"]
    assert r["fmax"] == pytest.approx(20.0, abs=1e-12)
    assert r["vapos"][:, 0] == pytest.approx([0.2, -0.2], abs=1e-14)
    assert r["apos"][:, 0] == pytest.approx([0.002, 0.998], abs=1e-14)
    steps, x, v = O.run_md(s, 1, 0.01, 1e-3, 50.0, 0.1)
    assert steps == 1
    assert v[:, 0] == pytest.approx([0.2, -0.2], abs=1e-14)
    assert x[:, 0] == pytest.approx([0.002, 0.998], abs=1e-14)
