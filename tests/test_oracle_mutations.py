"""Mutation self-test of the oracle pins.

Each entry below plants one plausible mistake (a flipped sign, a swapped
operand, a dropped term or clip, a wrong index or spacing) in a copy of
oracle/tsu_oracle.c, builds that copy with the oracle's own flags, and runs
tests/test_oracle.py against it (pyoracle loads the library named by
TSU_ORACLE_LIB).  Every mutant must make at least one pin fail: a pin suite
that a mistake survives does not pin that part of the oracle.  The
unmutated copy, built the same way, must pass.

All of this runs on the CPU; the mistakes are planted in the oracle only.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CFLAGS = ["-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# (name, reference line the oracle line restates, original text, mutated text)
MUTATIONS = [
    ("bed row upwinding reversed", "VolumeSolver.cu:110-115",
     "    if (s2 == s1)\n", "    if (s2 != s1)\n"),
    ("bed row always +a1 df3", "VolumeSolver.cu:110-115",
     "        vis[3] = -a1 * df[3];", "        vis[3] = a1 * df[3];"),
    ("depth-slope clip dropped", "Reconstructor.cu:150-152",
     "    if (0.5 * fabs(dh) > hi) dh = 2.0 * hi * o_sgn(dh);\n", ""),
    ("hc face slopes swapped", "Reconstructor.cu:191-194",
     "    Fm[4] = (ci - 0.5 * du) * Fm[0];\n    Fp[4] = (ci + 0.5 * du) * Fp[0];",
     "    Fm[4] = (ci + 0.5 * du) * Fm[0];\n    Fp[4] = (ci - 0.5 * du) * Fp[0];"),
    ("suspended row without a0 dU", "VolumeSolver.cu:98",
     "    DU[4] = hc1 - hc0;", "    DU[4] = 0.0;"),
    ("non-erodible mask ignored", "SedFlux.cuh:175",
     "    F[3] = o_bedload(p, s, k, h, ux, uy) * fixed * s->bedload;",
     "    F[3] = o_bedload(p, s, k, h, ux, uy) * s->bedload;"),
    ("van Leer operands transposed", "Reconstructor.cu:25",
     "        d = (a * bb + b * aa) / (aa + bb);", "        d = (a * aa + b * bb) / (aa + bb);"),
    ("dry threshold strict", "SedFlux.cuh:29",
     "    double x = qx / (h + O_EPSILON) * (h >= O_EPS);", "    double x = qx / (h + O_EPSILON) * (h > O_EPS);"),
    ("velocity cap without magnitude", "SedFlux.cuh:33-36",
     "        x = x / modu * O_VMAX;", "        x = x / modu;"),
    ("negative concentration kept", "SedFlux.cuh:123-125",
     "    c = c * (c > 0.0);\n", ""),
    ("Roe velocity weights swapped", "SedFlux.cuh:185-186",
     "    umx = (sqrt(h0) * u0x + sqrt(h1) * u1x)", "    umx = (sqrt(h0) * u1x + sqrt(h1) * u0x)"),
    ("S_L takes the larger speed", "VolumeSolver.cu:69-70",
     "    SL = fmin(SL, umx - ev_c);", "    SL = fmax(SL, umx - ev_c);"),
    ("a1 sign", "VolumeSolver.cu:77",
     "        a1 = (fabs(SR) - fabs(SL)) / (SR - SL);", "        a1 = (fabs(SL) - fabs(SR)) / (SR - SL);"),
    ("edge pressure term sign", "VolumeSolver.cu:92",
     "    df[1] += O_G * hm * deta;", "    df[1] -= O_G * hm * deta;"),
    ("hydrostatic reconstruction with max", "VolumeSolver.cu:38-40",
     "    zmH = fmin(zb0, zb1);", "    zmH = fmax(zb0, zb1);"),
    ("Dm carries F1", "VolumeSolver.cu:122",
     "        Dm[i] = 0.5 * (df[i] - vis[i]) + F0[i];", "        Dm[i] = 0.5 * (df[i] - vis[i]) + F1[i];"),
    ("rotation back sign", "VolumeSolver.cu:128-169",
     "    Dmo[1] = (normal_x * Dm[1] - normal_y * Dm[2]);", "    Dmo[1] = (normal_x * Dm[1] + normal_y * Dm[2]);"),
    ("friction from the stage state", "Reconstructor.cu:38-40",
     "    double h = S->V0[pos], ux, uy, modu, c;\n    orc_velocity(h, S->V0[1 * nv + pos], S->V0[2 * nv + pos], &ux, &uy);",
     "    double h = S->V[pos], ux, uy, modu, c;\n    orc_velocity(h, S->V[1 * nv + pos], S->V[2 * nv + pos], &ux, &uy);"),
    ("friction multiplies", "Reconstructor.cu:52",
     "    S->Vk[1 * nv + pos] /= (1.0 + c);", "    S->Vk[1 * nv + pos] *= (1.0 + c);"),
    ("Manning exponent 3/2", "(configuration law)",
     "(hp * cbrt(hp))", "(hp * sqrt(hp))"),
    ("exchange raises H", "Reconstructor.cu:86",
     "    S->Vk[3 * nv + pos] -= dt * phi;", "    S->Vk[3 * nv + pos] += dt * phi;"),
    ("exchange concentration with DBL_EPSILON", "Reconstructor.cu:71",
     "    c = hcv / (h + p->epsilon_h);", "    c = hcv / (h + O_EPSILON);"),
    ("bed face slope sign", "Reconstructor.cu:165-166",
     "    Fm[3] = Hi - 0.5 * dz;", "    Fm[3] = Hi + 0.5 * dz;"),
    ("bed slope as deta - dh", "Reconstructor.cu:164",
     "    dz = dh - deta;", "    dz = deta - dh;"),
    ("depth slope reversed", "Reconstructor.cu:140",
     "    dl = hi - hl;", "    dl = hl - hi;"),
    ("intra-cell pressure sign", "Reconstructor.cu:197",
     "                S->Vk[1 * nv + pos] -= dta * O_G * S->V[pos] * deta;",
     "                S->Vk[1 * nv + pos] += dta * O_G * S->V[pos] * deta;"),
    ("y source with dx", "Reconstructor.cu:274",
     "                dta = dt * inv_dy;", "                dta = dt * inv_dx;"),
    ("y fluxes scaled by dt/dx", "Volume_kernel.cu:191-195",
     "invy * (S->FD[f * nv + pos] + S->FU[f * nv + pos])", "invx * (S->FD[f * nv + pos] + S->FU[f * nv + pos])"),
    ("no RK averaging", "Volume_kernel.cu:245-253",
     "    double coef0 = 0.0 * (k == 0) + 0.5 * (k == 1);", "    double coef0 = 0.0;"),
    ("cell dt with dx on both axes", "Volume_kernel.cu:38-40",
     "    return p->CFL * 0.5 * fmin(p->dx / lx, p->dy / ly);", "    return p->CFL * 0.5 * fmin(p->dx / lx, p->dx / ly);"),
    ("dt not narrowed to float", "Reduction_kernel_min.cu:46-94",
     "    if (narrow && n > 4096) m = (double)(float)m;\n", ""),
    ("wall does not reflect", "VolumeSolver.cu:245-247",
     "    if (type == ORC_BC_WALL) g[ax] = -in[ax];", "    if (type == ORC_BC_WALL) g[ax] = in[ax];"),
    ("periodic neighbour clamped", "Reconstructor.cu:130-133",
     "        return periodic ? n - 1 : 0;", "        return 0;"),
    ("periodic seam flux to the wrong slot", "VolumeSolver.cu:217-408",
     "                    o_put5(S->FL, nv, q1, Dp);", "                    o_put5(S->FL, nv, q1, Dm);"),
    ("right inflow reads the bottom slots", "VolumeSolver.cu:172-202",
     "                    o_ghost(S, 1, Lf, G5, (long)npx + j, 1);", "                    o_ghost(S, 1, Lf, G5, (long)j, 1);"),
    ("bottom ghost negates qx", "VolumeSolver.cu:285-287",
     "                o_ghost(S, 2, Lf, G5, (long)i, 2);", "                o_ghost(S, 2, Lf, G5, (long)i, 1);"),
    ("inflow depth without H", "VolumeSolver.cu:196",
     "    h1 += zb1;\n", ""),
    ("hmax on dry cells", "Volume_kernel.cu:86-88",
     "        if (h > eh) S->hmax[pos] = fmax(S->hmax[pos], h);", "        S->hmax[pos] = fmax(S->hmax[pos], h);"),
    ("Grass law |u| instead of |u|^2", "(configuration law)",
     "        qb = k->grass_k * u * modu * modu", "        qb = k->grass_k * u * modu"),
    ("bedload direction by |u|", "SedFlux.cuh:163",
     "    K1 = 17.0 * u / (modu + O_EPSILON);", "    K1 = 17.0 * modu / (modu + O_EPSILON);"),
    ("settling velocity with R", "Reconstructor.cu:64-66",
     "+ 1.09 * sd * O_G * p->DS)", "+ 1.09 * p->R * O_G * p->DS)"),
    ("near-bed ratio dropped", "SedFlux.cuh:83-85",
     "    Cb = k->cbk * c;", "    Cb = c;"),
    ("entrainment Z^4", "SedFlux.cuh:66-68",
     "    Z5 = Z * Z * Z * Z * Z;", "    Z5 = Z * Z * Z * Z;"),
]


def _build(tmp, name, original, mutated):
    src = open(os.path.join(ORACLE, "tsu_oracle.c")).read()
    if original is not None:
        n = src.count(original)
        if n != 1:
            raise AssertionError("mutation %r: the original text occurs %d times" % (name, n))
        src = src.replace(original, mutated)
    d = os.path.join(tmp, "m%08x" % (abs(hash(name)) & 0xffffffff))
    os.makedirs(d)
    for f in ("tsu_oracle.h", "tsu_quotient.c"):
        shutil.copy(os.path.join(ORACLE, f), d)
    with open(os.path.join(d, "tsu_oracle.c"), "w") as fh:
        fh.write(src)
    so = os.path.join(d, "libmut.so")
    subprocess.check_call(["gcc"] + CFLAGS + ["-o", so, os.path.join(d, "tsu_oracle.c"),
                                              os.path.join(d, "tsu_quotient.c"), "-lm"])
    return so


def _run_pins(so):
    env = dict(os.environ, TSU_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-x", "-q",
                        "-p", "no:cacheprovider"], env=env, cwd=ROOT, capture_output=True, text=True, timeout=600)
    failed = [ln.split(" ")[1] for ln in r.stdout.splitlines() if ln.startswith("FAILED ")]
    return r.returncode, failed, r.stdout[-2000:]


def test_unmutated_copy_passes(tmp_path):
    rc, failed, out = _run_pins(_build(str(tmp_path), "unmutated", None, None))
    assert rc == 0, out


def test_every_mutation_fails_a_pin(tmp_path):
    sos = [(name, ref, _build(str(tmp_path), name, a, b)) for name, ref, a, b in MUTATIONS]
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda t: (t[0], t[1]) + _run_pins(t[2]), sos))
    survivors = []
    for name, ref, rc, failed, out in results:
        print("%-40s %-28s rc=%d caught by %s" % (name, ref, rc, failed[:1]))
        # rc 1: tests ran and some failed (a crash or a collection error is not a pin failing)
        if rc != 1 or not failed:
            survivors.append((name, ref, rc, out[-600:]))
    assert not survivors, "mutations no pin catches:\n" + "\n".join("%s (%s) rc=%d\n%s" % s for s in survivors)
