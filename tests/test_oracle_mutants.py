"""The oracle's pins catch plausible mistakes.

Each case plants one mistake in a copy of oracle/fs_oracle.cpp (a flipped
sign, a dropped term, a transposed operand, a swapped weight, a wrong
staggering offset), builds that copy into a temporary library and runs the
pins of tests/test_oracle_pins.py (and test_oracle_face_fractions.py) that
cover the function against it: at
least one of them must fail with an assertion.  The copies live in pytest's
temporary directory; the oracle itself is not modified.
"""
import os
import shutil
import subprocess

import pytest

import test_oracle_face_fractions as ff
import test_oracle_pins as pins
from fluidprjct_amd._abi import FS_STEP_FACE_FRACTIONS, FsLib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
FLAGS = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-w"]

# (name, text in fs_oracle.cpp (exactly once), replacement, pins that must catch it)
MUTANTS = [
    ("lift sign",
     "Fl = -0.5f * rho * Cl * Aeff * ura * cross(urel, nxu / nxa);",
     "Fl = 0.5f * rho * Cl * Aeff * ura * cross(urel, nxu / nxa);",
     [lambda lib: pins.test_drag_lift_one_face_closed_form(lib, "+x", (1.0, 1.0, 0.0))]),
    ("drag dropped",
     "f3 Ft = Fd + Fl, Tt = Td + Tl;",
     "f3 Ft = Fl, Tt = Tl;",
     [lambda lib: pins.test_drag_lift_one_face_closed_form(lib, "-z", (0.3, -0.4, -1.2))]),
    ("effective area without the normal term",
     "if (nur > 0) Aeff = (nur / ura * alpha_s + (1 - alpha_s)) * area;",
     "if (nur > 0) Aeff = area;",
     [lambda lib: pins.test_drag_lift_one_face_closed_form(lib, "+x", (1.0, 1.0, 0.0))]),
    ("torque taken in the world frame",
     "f3 tau_local = mk(R[0] * torque.x + R[3] * torque.y + R[6] * torque.z,\n"
     "                          R[1] * torque.x + R[4] * torque.y + R[7] * torque.z,\n"
     "                          R[2] * torque.x + R[5] * torque.y + R[8] * torque.z);",
     "f3 tau_local = mk(R[0] * torque.x + R[1] * torque.y + R[2] * torque.z,\n"
     "                          R[3] * torque.x + R[4] * torque.y + R[5] * torque.z,\n"
     "                          R[6] * torque.x + R[7] * torque.y + R[8] * torque.z);",
     [lambda lib: pins.test_integrator_torque_in_body_frame(lib)]),
    ("Euler rate sign",
     "float dphi_dt = (float)(omega.x + dpsi_dt * std::sin((double)theta));",
     "float dphi_dt = (float)(omega.x - dpsi_dt * std::sin((double)theta));",
     [lambda lib: pins.test_euler_angle_rates(lib)]),
    ("rotation matrix entry",
     "R[3] = (float)(std::sin(ps));",
     "R[3] = (float)(-std::sin(ps));",
     [lambda lib: pins.test_rotation_matrix_is_a_rotation(lib),
      lambda lib: pins.test_integrator_torque_in_body_frame(lib)]),
    ("P2G weights swapped along x",
     "float wx = (dxi == 0 ? (1.0f - b.fx) : b.fx);",
     "float wx = (dxi == 0 ? b.fx : (1.0f - b.fx));",
     [lambda lib: pins.test_p2g_single_particle_weights(lib), lambda lib: pins.test_p2g_two_particles_hat_weights(lib)]),
    ("u sampled without the half-cell offset in y",
     "b.i = (int)floorf(rx); b.j = (int)floorf(ry - 0.5f); b.k = (int)floorf(rz - 0.5f);",
     "b.i = (int)floorf(rx); b.j = (int)floorf(ry); b.k = (int)floorf(rz - 0.5f);",
     [lambda lib, a=a: pins.test_g2p_samples_faces_exactly(lib, a) for a in (None, 0, 1, 2)]),
    ("trilinear lerp weights swapped",
     "float c00 = fmaf(c000, 1.0f - xd, c100 * xd);",
     "float c00 = fmaf(c000, xd, c100 * (1.0f - xd));",
     [lambda lib, a=a: pins.test_g2p_samples_faces_exactly(lib, a) for a in (None, 0, 1, 2)]),
    ("gravity on every component",
     "const float acc[3] = {0.0f, s.cfg.gravity_y, 0.0f};",
     "const float acc[3] = {s.cfg.gravity_y, s.cfg.gravity_y, 0.0f};",
     [lambda lib: pins.test_hydrostatic_pool(lib)]),
    ("divergence sign",
     "float rhs_val = -scale * div;",
     "float rhs_val = scale * div;",
     [lambda lib: pins.test_hydrostatic_pool(lib), lambda lib: pins.test_rhs_of_a_linear_field(lib)]),
    ("solid face term sign",
     "rhs_val -= scale * (u[ui0] - us_);",
     "rhs_val += scale * (u[ui0] - us_);",
     [lambda lib: pins.test_hydrostatic_pool(lib), lambda lib: pins.test_rhs_of_a_linear_field(lib)]),
    ("body face at the velocity of the wrong axis",
     "rhs_val += scale * (v[vi0 + W] - vs_);",
     "rhs_val += scale * (v[vi0 + W] - (L == FS_BODY ? body_vel(fp).x : 0.0f));",
     [lambda lib: pins.test_rhs_of_a_linear_field(lib)]),
    ("pressure gradient sign",
     "G[idx] = g - scale * (p1 - p0);",
     "G[idx] = g + scale * (p1 - p0);",
     [lambda lib: pins.test_hydrostatic_pool(lib)]),
    ("AIR neighbour left out of the diagonal",
     "            diag += 1.0;\n            fdiag = fdiag + th;",
     "            if (L[nidx] == FS_FLUID) diag += 1.0;\n            fdiag = fdiag + th;",
     [lambda lib: pins.test_reference_csr_of_enclosed_toy(lib), lambda lib: pins.test_hydrostatic_pool(lib),
      lambda lib: pins.test_direct_solve_vs_scipy(lib, "dam")]),
    ("buoyancy sign",
     "f3 Farch = (-rho * subf) * G;",
     "f3 Farch = (rho * subf) * G;",
     [lambda lib: pins.test_body_sphere_labels_and_buoyancy(lib)]),
    ("RK3 second stage at the wrong fraction",
     "f3 k2 = interp_vel(F, pos + k1 * 0.5f * dT);",
     "f3 k2 = interp_vel(F, pos + k1 * 0.75f * dT);",
     [lambda lib: pins.test_advection_rk3_in_a_rotation_field(lib)]),
    ("RK3 weights swapped",
     "pos = pos + (2.0f / 9.0f) * dT * k1 + (3.0f / 9.0f) * dT * k2 + (4.0f / 9.0f) * dT * k3;",
     "pos = pos + (2.0f / 9.0f) * dT * k1 + (4.0f / 9.0f) * dT * k2 + (3.0f / 9.0f) * dT * k3;",
     [lambda lib: pins.test_advection_rk3_in_a_rotation_field(lib)]),
    ("seeding octant offset",
     "const float off = 0.25f * dx;",
     "const float off = 0.3f * dx;",
     [lambda lib: pins.test_seeding_matches_reference_init_dump(lib)]),
    ("FLIP blend weights swapped",
     "float nu = alpha * pic.x + (1.0f - alpha) * (s.vel[3 * q] + flip.x);",
     "float nu = (1.0f - alpha) * pic.x + alpha * (s.vel[3 * q] + flip.x);",
     [lambda lib: pins.test_g2p_uniform_fields_and_advection(lib)]),
    ("triangle fraction with one edge ratio twice",
     "const float part = (p / (p - q)) * (p / (p - r));",
     "const float part = (p / (p - q)) * (p / (p - q));",
     [lambda lib: ff.test_open_fraction_of_a_plane_cut(lib)]),
    ("SDF lerp weights swapped along y",
     "const float c0 = c00 * (1.0f - fy) + c10 * fy, c1 = c01 * (1.0f - fy) + c11 * fy;",
     "const float c0 = c00 * fy + c10 * (1.0f - fy), c1 = c01 * fy + c11 * (1.0f - fy);",
     [lambda lib: ff.test_open_fraction_of_a_plane_cut(lib)]),
    ("cut face without the body velocity",
     "rhs_val -= scale * ((1.0f - th) * (u[ui0] - body_vel(fp).x));",
     "rhs_val -= scale * ((1.0f - th) * u[ui0]);",
     [lambda lib: ff.test_projection_makes_the_weighted_divergence_vanish(lib, FS_STEP_FACE_FRACTIONS)]),
    ("open fraction in place of the closed one",
     "rhs_val += scale * ((1.0f - th) * (v[vi0 + W] - body_vel(fp).y));",
     "rhs_val += scale * (th * (v[vi0 + W] - body_vel(fp).y));",
     [lambda lib: ff.test_projection_makes_the_weighted_divergence_vanish(lib, FS_STEP_FACE_FRACTIONS)]),
    ("unweighted diagonal with weighted neighbours",
     "row.push_back({r, open ? (double)fdiag : diag});",
     "row.push_back({r, diag});",
     [lambda lib: ff.test_hydrostatic_pool_with_a_resting_body(lib),
      lambda lib: ff.test_pressure_matches_scipy_on_the_weighted_matrix(lib)]),
    ("closed face keeps its velocity",
     "if (ff && face_open_sdf(s, comp, fi, fj, fk) == 0.0f) p1 = p0 + invScale * (g - vbc);",
     "if (ff && face_open_sdf(s, comp, fi, fj, fk) < 0.0f) p1 = p0 + invScale * (g - vbc);",
     [lambda lib: ff.test_thin_plate_closes_the_faces_it_covers(lib)]),
]


@pytest.fixture(scope="module")
def lnet_object(tmp_path_factory):
    d = tmp_path_factory.mktemp("lnet")
    obj = str(d / "lnet_oracle.o")
    subprocess.run(["g++", *FLAGS, "-c", os.path.join(ORACLE, "lnet_oracle.cpp"), "-o", obj], check=True)
    return obj


def build_mutant(tmp_path, old, new, lnet_object):
    src = open(os.path.join(ORACLE, "fs_oracle.cpp")).read()
    assert src.count(old) == 1, f"mutation site not unique or missing: {old!r}"
    (tmp_path / "oracle").mkdir()
    shutil.copytree(os.path.join(ROOT, "include"), tmp_path / "include")
    (tmp_path / "oracle" / "fs_oracle.cpp").write_text(src.replace(old, new))
    obj = str(tmp_path / "fs_oracle.o")
    lib = str(tmp_path / "libfs_mutant.so")
    subprocess.run(["g++", *FLAGS, "-c", str(tmp_path / "oracle" / "fs_oracle.cpp"), "-o", obj], check=True)
    subprocess.run(["g++", "-shared", "-fopenmp", "-o", lib, obj, lnet_object], check=True)
    return lib


@pytest.mark.parametrize("name,old,new,checks", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(name, old, new, checks, tmp_path, lnet_object):
    lib = FsLib(build_mutant(tmp_path, old, new, lnet_object))
    caught = []
    for check in checks:
        try:
            check(lib)
        except AssertionError:
            caught.append(True)
            break
    assert caught, f"no pin fails with the planted mistake '{name}'"
