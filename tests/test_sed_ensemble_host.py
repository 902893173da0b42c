"""The vectorised particle bank of settling ensembles against the per-particle code that the
single-run GPU tests pin to the reference fixtures: bit for bit, over mixed members (one and
two particles, DF / MDF / DFC, marker rings of different sizes), with random marker forces
standing in for the device.  CPU only."""
import dataclasses
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import golden, golden_cfg
from iblbm_amd.init import initialize
from iblbm_amd.sed_ensemble import ParticleBank, _trig_vector_matches_scalar
from iblbm_amd.sedimentation import (full_step_particles, half_step_particles, make_particles,
                                     marker_forces_to_particles)

FIXTURES = ["sed_single_df_bgk_hat", "sed_single_mdf_cm_p4", "sed_multi_dfc_cm_hat",
            "sed_multi_df_bgk_hat", "sed_multi_mdf_cm_p4"]


def _member(cfg):
    st = initialize(cfg, fields=False)
    return SimpleNamespace(cfg=cfg, st=st, particles=make_particles(cfg, st), g=cfg.gravity,
                           multi=cfg.particles_config is not None)


def _members(variant):
    out = []
    for name in FIXTURES:
        cfg = golden_cfg(golden(name)[1])
        if variant == "semi_implicit":
            cfg = dataclasses.replace(cfg, rotation_coupling="semi_implicit")
        elif variant == "no_rotation":
            cfg = dataclasses.replace(cfg, enable_rotation=False)
        out.append(_member(cfg))
    return out


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


@pytest.mark.parametrize("variant", ["indirect", "semi_implicit", "no_rotation"])
def test_bank_matches_per_particle_code(variant):
    scalar, vec = _members(variant), _members(variant)
    bank = ParticleBank(vec)
    rng = np.random.default_rng(11)
    for step in range(40):
        Lx, Ly, ud = bank.half_step()
        for m in scalar:
            half_step_particles(m.particles, m.cfg, m.st, m.g, m.multi)
        parts = [p for m in scalar for p in m.particles]
        assert np.array_equal(_bits(Lx), _bits(np.concatenate([p.Lx for p in parts])))
        assert np.array_equal(_bits(Ly), _bits(np.concatenate([p.Ly for p in parts])))
        assert np.array_equal(_bits(ud), _bits(np.concatenate([p.desired for p in parts])))
        q = rng.normal(scale=1e-3, size=(bank.n_markers, 2))
        F, T = bank.forces_from_markers(q)
        c, k = 0, 0
        for m in scalar:
            lbt = sum(p.Lb for p in m.particles)
            out = marker_forces_to_particles(m.particles, q[c:c + lbt], m.cfg, m.st)
            c += lbt
            for f, t in out:
                assert np.array_equal(_bits(F[k]), _bits(f)), (step, k)
                assert _bits(T[k]) == _bits(t), (step, k)
                k += 1
            full_step_particles(m.particles, out, m.cfg, m.st, m.g)
        bank.full_step(F, T)
        for name in ["pos", "vel", "vel_prev", "force"]:
            assert np.array_equal(_bits(getattr(bank, name)),
                                  _bits([getattr(p, name) for p in parts])), (step, name)
        for name in ["omega", "angle", "torque"]:
            assert np.array_equal(_bits(getattr(bank, name)),
                                  _bits([float(getattr(p, name)) for p in parts])), (step, name)


@pytest.mark.parametrize("scheme", ["new_velocity", "trapezoidal"])
def test_bank_euler_matches_per_particle_code(scheme):
    """The vectorised explicit Euler update against ``euler_step_particles``, with random
    forces and torques and a random rho / U field standing in for the lattice (each box read
    from it): positions, velocities, rotation, the internal angular momentum and the next
    step's rings and desired velocities, bit for bit."""
    from iblbm_amd.sedimentation import euler_step_particles

    def members():
        out = []
        for name in FIXTURES:
            cfg = dataclasses.replace(golden_cfg(golden(name)[1]),
                                      time_integrator="euler_explicit",
                                      sedimentation_euler_update_scheme=scheme)
            out.append(_member(cfg))
        return out

    scalar, vec = members(), members()
    bank = ParticleBank(vec)
    assert bank.euler
    rng = np.random.default_rng(12)
    fields = {}

    def field(mi, m):  # a fixed rho / U field per member and step
        if mi not in fields:
            nx, ny = m.st.nx, m.st.ny
            fields[mi] = (1.0 + 0.01 * rng.standard_normal((ny, nx)),
                          0.01 * rng.standard_normal((ny, nx, 2)))
        return fields[mi]

    owner = [mi for mi, m in enumerate(scalar) for _ in m.particles]
    for step in range(30):
        fields.clear()
        F = rng.normal(scale=1e-3, size=(bank.P, 2))
        T = rng.normal(scale=1e-4, size=bank.P)
        Lx, Ly, ud = bank.euler_step(F, T, lambda k, x0, y0, w, h: tuple(
            a[y0:y0 + h, x0:x0 + w] for a in field(owner[k], scalar[owner[k]])))
        k = 0
        for mi, m in enumerate(scalar):
            n = len(m.particles)
            euler_step_particles(m.particles, [(F[k + j], T[k + j]) for j in range(n)], m.cfg,
                                 m.st, m.g, lambda x0, y0, w, h, mi=mi, m=m: tuple(
                                     a[y0:y0 + h, x0:x0 + w] for a in field(mi, m)))
            k += n
        parts = [p for m in scalar for p in m.particles]
        for name in ["pos", "vel", "vel_prev", "force"]:
            assert np.array_equal(_bits(getattr(bank, name)),
                                  _bits([getattr(p, name) for p in parts])), (step, name)
        for name in ["omega", "angle", "torque", "l_int_prev"]:
            assert np.array_equal(_bits(getattr(bank, name)),
                                  _bits([float(getattr(p, name)) for p in parts])), (step, name)
        assert np.array_equal(_bits(Lx), _bits(np.concatenate([p.Lx for p in parts])))
        assert np.array_equal(_bits(Ly), _bits(np.concatenate([p.Ly for p in parts])))
        assert np.array_equal(_bits(ud), _bits(np.concatenate([p.desired for p in parts])))


def test_frozen_member_keeps_its_particles():
    members = _members("indirect")
    bank = ParticleBank(members)
    bank.live[:1] = False  # the first member (one particle) stopped
    pos0, vel0, ang0 = bank.pos[0].copy(), bank.vel[0].copy(), bank.angle[0]
    bank.half_step()
    bank.full_step(np.ones((bank.P, 2)) * 1e-3, np.ones(bank.P) * 1e-4)
    assert np.array_equal(bank.pos[0], pos0) and np.array_equal(bank.vel[0], vel0)
    assert bank.angle[0] == ang0
    assert not np.array_equal(bank.pos[1], members[1].particles[0].pos)


def test_vector_trig_check_runs():
    assert _trig_vector_matches_scalar() in (True, False)


def test_two_disc_cases():
    from iblbm_amd.sed_ensemble import two_disc_cfgs
    cfgs = two_disc_cfgs(64, seed=5)
    d = 1.0 / 8
    gaps, offs = [], []
    for cfg in cfgs:
        m = _member(cfg)
        assert (m.st.nx, m.st.ny) == (193, 1921)
        assert [p.Lb for p in m.particles] == [75, 75]
        assert abs(2.0 * m.st.lattice_r - 24.0) < 1e-12
        (x0, y0), (x1, y1) = (pc["center"] for pc in cfg.particles_config)
        gaps.append((y1 - y0) / d)
        offs.append((x1 - x0) / d)
        assert {pc["rho_ratio"] for pc in cfg.particles_config} == {1.01}
        assert y1 + d / 2 < cfg.ymax - 2.0 * m.st.dx
    assert 2.0 <= min(gaps) and max(gaps) <= 4.0 and max(gaps) - min(gaps) > 1.0
    assert 0.0 <= min(offs) and max(offs) <= 0.1
    assert len({round(g, 12) for g in gaps}) == 64
    assert two_disc_cfgs(3, seed=5)[2] == cfgs[2]
