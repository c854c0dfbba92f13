"""Generate the golden fixtures under tests/golden/ from the reference itself.

Runs only where the reference was compiled from its sources
(`make -C oracle ref` -> oracle/_ref/cbns3d_ref, oracle/_ref/libref_point.so,
both built from the reference's own files with the image's MPICH); the GPU
box never runs this.  Each fixture is data: inputs and the reference's
outputs.

Fixtures
  ref_tgv16.npz      TGV 16^3, interior Q after 1, 2 and 10 steps (-np 1);
                     -np 8 (2 2 2) equality flag
  ref_uneven.npz     TGV 26x20x16, 3 steps, -np 6 (3 2 1) with the remainder
                     partition, assembled interior Q; -np 1 equality flag
  ref_tgv64.npz      TGV 64^3 digests after 1 and 100 steps (-np 8, 2 2 2)
  ref_restart16.npz  checkpoint-injected state (is_restart = 0) under
                     MALLOC_PERTURB_=255, 5 steps, run twice
  ref_turb16.npz     seeded decaying turbulence (M_t = 0.6, Re_lambda = 100,
                     cbns3d_cart_mpi_amd.turbulence, seed 7) at 16^3 injected
                     through the checkpoint (is_restart = 0), 10 steps
  ref_vtk8.npz       TGV 8^3, 2 steps: the reference's VTK files (output000,
                     output001) and checkpoint, as bytes
  ref_point.npz      the reference's compiled weno5 / rec_weno5 /
                     calc_eigenvectors / lf_flux on seeded inputs
"""
import ctypes
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

MPIEXEC = "/opt/conda/bin/mpiexec"
NG = 4


def input_txt(N, nsteps, is_restart=1, t_end=20.0, mach="0.1", re="1600.0"):
    # the 17 positional values of src/read_input.cpp:24-72, python/taylor_green.ini values
    vals = ["checkpoint", "0.5", "0.0", repr(t_end), "0.71", "1.4", mach, re, "255.556", "1.0",
            str(N[0]), str(N[1]), str(N[2]), str(nsteps), str(nsteps), "1", str(is_restart)]
    return "\n".join(vals) + "\n"


def run_ref(N, nsteps, dims=(1, 1, 1), is_restart=1, pre=None, env_extra=None, **kw):
    d = tempfile.mkdtemp(prefix="ref_", dir=os.environ.get("TMPDIR", "/tmp"))
    with open(os.path.join(d, "input.txt"), "w") as f:
        f.write(input_txt(N, nsteps, is_restart, **kw))
    if pre:
        pre(d)
    env = dict(os.environ)
    if env_extra:
        env.update(env_extra)
    nproc = dims[0] * dims[1] * dims[2]
    # src/cbns3d.cpp:80 broadcasts SimulationParams (which holds a std::string)
    # as raw bytes; the string's pointer is only valid on other ranks when
    # every process has the same address layout, so multi-rank runs go
    # without address-space randomisation.
    pre_cmd = ["setarch", "x86_64", "-R"] if nproc > 1 else []
    out = subprocess.run(pre_cmd + [MPIEXEC, "-np", str(nproc), O.REF_BIN_PATH] + [str(x) for x in dims],
                         cwd=d, env=env, check=True, capture_output=True, text=True)
    return d, out.stdout


def read_ckpt(fn):
    """src/Block3d.cpp:964-987: [f64 t][u32 IM_G][u32 JM_G][u32 KM_G][f64 Q AoS]"""
    b = open(fn, "rb").read()
    t = float(np.frombuffer(b[:8], np.float64)[0])
    IG, JG, KG = (int(x) for x in np.frombuffer(b[8:20], np.uint32))
    Q = np.frombuffer(b[20:], np.float64).reshape(KG, JG, IG, 5).copy()
    return t, Q


def write_ckpt(fn, t, Qg):
    KG, JG, IG, _ = Qg.shape
    with open(fn, "wb") as f:
        f.write(np.float64(t).tobytes())
        f.write(np.array([IG, JG, KG], np.uint32).tobytes())
        f.write(np.ascontiguousarray(Qg, np.float64).tobytes())


def assemble(d, N, dims):
    """Interior of the global domain from the per-rank checkpoints.  Ranks map
    to Cartesian coordinates row-major (last dimension fastest); blocks hold
    N/dims cells except the last, which takes the remainder (src/cbns3d.cpp:102-126)."""
    out = np.empty((N[2], N[1], N[0], 5))
    t0 = None
    r = 0
    for cx in range(dims[0]):
        for cy in range(dims[1]):
            for cz in range(dims[2]):
                t, Q = read_ckpt(os.path.join(d, "checkpoint%03d.dat" % r))
                t0 = t if t0 is None else t0
                assert t == t0
                x0, y0, z0 = cx * (N[0] // dims[0]), cy * (N[1] // dims[1]), cz * (N[2] // dims[2])
                KG, JG, IG, _ = Q.shape
                out[z0:z0 + KG - 2 * NG, y0:y0 + JG - 2 * NG, x0:x0 + IG - 2 * NG] = Q[NG:-NG, NG:-NG, NG:-NG]
                r += 1
    return t0, out


def digest(Q):
    """Order-fixed digest of an interior AoS state (signed zeros normalised)."""
    Qn = np.ascontiguousarray(Q + 0.0)
    f = Qn.reshape(-1, 5)
    return dict(
        sha256=np.frombuffer(hashlib.sha256(Qn.tobytes()).digest(), np.uint8),
        sum=np.array([np.sum(f[:, e], dtype=np.float64) for e in range(5)]),
        linf=np.abs(f).max(axis=0),
        l2=np.sqrt((f * f).sum(axis=0)),
    )


def sample(Q, n=4096, seed=1234):
    rng = np.random.default_rng(seed)
    flat = Q.reshape(-1, 5)
    idx = np.sort(rng.choice(flat.shape[0], size=n, replace=False))
    return idx, flat[idx].copy()


def gen_tgv16():
    N = (16, 16, 16)
    res = {}
    for n in (1, 2, 10):
        d, _ = run_ref(N, n)
        t, Q = read_ckpt(os.path.join(d, "checkpoint000.dat"))
        res["Q_n%d" % n] = Q[NG:-NG, NG:-NG, NG:-NG].copy()
        res["t_n%d" % n] = np.float64(t)
        shutil.rmtree(d)
    d, _ = run_ref(N, 2, dims=(2, 2, 2))
    t8, Q8 = assemble(d, N, (2, 2, 2))
    shutil.rmtree(d)
    res["np8_equals_np1_n2"] = np.bool_(np.array_equal(Q8, res["Q_n2"]) and t8 == res["t_n2"])
    np.savez_compressed(os.path.join(HERE, "ref_tgv16.npz"), **res)
    print("ref_tgv16: np8==np1:", bool(res["np8_equals_np1_n2"]))


def gen_uneven():
    N, dims, n = (26, 20, 16), (3, 2, 1), 3
    d, _ = run_ref(N, n, dims=dims)
    t, Q = assemble(d, N, dims)
    shutil.rmtree(d)
    d, _ = run_ref(N, n)
    t1, Q1 = read_ckpt(os.path.join(d, "checkpoint000.dat"))
    shutil.rmtree(d)
    eq = np.array_equal(Q, Q1[NG:-NG, NG:-NG, NG:-NG]) and t == t1
    np.savez_compressed(os.path.join(HERE, "ref_uneven.npz"), Q=Q, t=np.float64(t), N=np.array(N),
                        dims=np.array(dims), nsteps=np.int64(n), np6_equals_np1=np.bool_(eq))
    print("ref_uneven: np6==np1:", eq)


def gen_tgv64():
    N = (64, 64, 64)
    res = {}
    for n in (1, 100):
        d, out = run_ref(N, n, dims=(2, 2, 2))
        t, Q = assemble(d, N, (2, 2, 2))
        shutil.rmtree(d)
        res["t_n%d" % n] = np.float64(t)
        for k, v in digest(Q).items():
            res["%s_n%d" % (k, n)] = v
        idx, vals = sample(Q)
        res["sample_idx_n%d" % n] = idx
        res["sample_val_n%d" % n] = vals
        if n == 100:
            # per-step dt as printed by the driver (6 significant digits, src/cbns3d.cpp:196-198)
            dts = [float(line.split(",")[1]) for line in out.splitlines()
                   if line.count(",") == 2 and line.split(",")[0].strip().isdigit()]
            res["dt_printed"] = np.array(dts)
    np.savez_compressed(os.path.join(HERE, "ref_tgv64.npz"), **res)
    print("ref_tgv64 done")


def gen_restart16():
    N = (16, 16, 16)
    p = O.make_params(N)
    b = O.Block(p)
    b.init_tgv()
    Qg = b.Q_ghosted(0)
    rng = np.random.default_rng(20240601)
    pert = rng.standard_normal(Qg[NG:-NG, NG:-NG, NG:-NG].shape) * np.array([1e-2, 1e-2, 1e-2, 1e-2, 1e-1])
    Qin = Qg.copy()
    Qin[NG:-NG, NG:-NG, NG:-NG] += pert
    t_in = 0.5
    outs = []
    for _ in range(2):
        d, _ = run_ref(N, 5, is_restart=0,
                       pre=lambda dd: write_ckpt(os.path.join(dd, "checkpoint000.dat"), t_in, Qin),
                       env_extra={"MALLOC_PERTURB_": "255"})
        outs.append(read_ckpt(os.path.join(d, "checkpoint000.dat")))
        shutil.rmtree(d)
    assert outs[0][0] == outs[1][0] and np.array_equal(outs[0][1], outs[1][1]), "restart run not deterministic"
    np.savez_compressed(os.path.join(HERE, "ref_restart16.npz"), Q_in=Qin[NG:-NG, NG:-NG, NG:-NG].copy(),
                        t_in=np.float64(t_in), Q_out=outs[0][1][NG:-NG, NG:-NG, NG:-NG].copy(),
                        t_out=np.float64(outs[0][0]), nsteps=np.int64(5))
    print("ref_restart16 done")


def gen_point():
    R = O.ref_point_lib()
    assert R is not None, "build oracle/_ref first (make -C oracle ref)"
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(777)
    gamma = 1.4
    # weno5: random, constant, linear, quadratic, jumps, tiny and huge scales
    W = [rng.standard_normal(5) * 10.0 ** rng.integers(-3, 3) + rng.standard_normal() * 100 for _ in range(1000)]
    W += [np.full(5, v) for v in (0.0, 1.0, -2.5, 71.42857142857143)]
    W += [a + b * np.arange(5.0) for a, b in rng.standard_normal((100, 2)) * 10]
    W += [a + b * np.arange(5.0) ** 2 for a, b in rng.standard_normal((100, 2))]
    W += [np.where(np.arange(5) < s, a, b) for s, a, b in zip(rng.integers(1, 5, 100), rng.standard_normal(100),
                                                               rng.standard_normal(100) * 10)]
    W = np.array(W)
    wout = np.array([R.refpt_weno5(*map(float, w)) for w in W])

    # physical states around the TGV / turbulence regimes
    def prim(n):
        rho = np.exp(rng.uniform(-1.0, 1.0, n))
        vel = rng.standard_normal((n, 3)) * rng.choice([0.01, 0.3, 1.0, 3.0], size=(n, 1))
        p = np.exp(rng.uniform(-1.0, 4.5, n))
        return np.column_stack([rho, vel, p])

    M = 450
    pL, pR = prim(M), prim(M)
    pR[:50] = pL[:50]  # equal states: lf_flux(U, U) = F(U)
    normals = np.zeros((M, 3))
    normals[np.arange(M), np.arange(M) % 3] = 1.0
    normals[-90:] = rng.standard_normal((90, 3))  # general normals exercise both frame branches
    normals[-90:-80] = [0.0, 0.0, 1.0]
    Rm = np.empty((M, 25))
    Lm = np.empty((M, 25))
    F = np.empty((M, 5))
    for m in range(M):
        a, b, nn = np.ascontiguousarray(pL[m]), np.ascontiguousarray(pR[m]), np.ascontiguousarray(normals[m])
        r, l_, f = np.empty(25), np.empty(25), np.empty(5)
        R.refpt_calc_eigenvectors(gamma, a.ctypes.data_as(P), b.ctypes.data_as(P), nn.ctypes.data_as(P),
                                  r.ctypes.data_as(P), l_.ctypes.data_as(P))
        R.refpt_lf_flux(gamma, b.ctypes.data_as(P), a.ctypes.data_as(P), nn.ctypes.data_as(P), f.ctypes.data_as(P))
        Rm[m], Lm[m], F[m] = r, l_, f
    # rec_weno5 on 5-cell stencils of conserved states, with the Roe frame of cells 2, 3
    K = 450
    stens = np.empty((K, 5, 5))
    for kk in range(K):
        pr = prim(5)
        if kk % 3 == 0:
            pr = pr[0] + 0.05 * (pr - pr[0])  # smooth stencils
        rho, u, v, w, p = pr.T
        stens[kk] = np.column_stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
    rec_out = np.empty((K, 5))
    for kk in range(K):
        f = np.ascontiguousarray(stens[kk].reshape(25))
        r = np.ascontiguousarray(Rm[kk % M])
        l_ = np.ascontiguousarray(Lm[kk % M])
        o = np.empty(5)
        R.refpt_rec_weno5(f.ctypes.data_as(P), r.ctypes.data_as(P), l_.ctypes.data_as(P), 0.4, o.ctypes.data_as(P))
        rec_out[kk] = o
    np.savez_compressed(os.path.join(HERE, "ref_point.npz"), weno_in=W, weno_out=wout, primL=pL, primR=pR,
                        normal=normals, R=Rm, L=Lm, flux=F, rec_in=stens, rec_out=rec_out, gamma=gamma, gam1=0.4)
    print("ref_point done")


def gen_turb16():
    from cbns3d_cart_mpi_amd.turbulence import turbulence_state

    N, seed, n = 16, 7, 10
    Q, Re = turbulence_state(N, seed=seed)
    Qg = np.zeros((N + 2 * NG,) * 3 + (5,))
    Qg[NG:-NG, NG:-NG, NG:-NG] = Q
    d, _ = run_ref((N, N, N), n, is_restart=0, mach="1.0", re=repr(Re),
                   pre=lambda dd: write_ckpt(os.path.join(dd, "checkpoint000.dat"), 0.0, Qg),
                   env_extra={"MALLOC_PERTURB_": "255"})
    t, Qo = read_ckpt(os.path.join(d, "checkpoint000.dat"))
    shutil.rmtree(d)
    np.savez_compressed(os.path.join(HERE, "ref_turb16.npz"), Q_in=Q, Re=np.float64(Re), Mach=np.float64(1.0),
                        seed=np.int64(seed), nsteps=np.int64(n), t_out=np.float64(t),
                        Q_out=Qo[NG:-NG, NG:-NG, NG:-NG].copy())
    print("ref_turb16 done, Re =", Re, "t =", t)


def gen_vtk8():
    N = (8, 8, 8)
    d, out = run_ref(N, 2)
    res = {}
    for nm in ("output000/output000.vtr", "output000/output.vtm", "output001/output000.vtr",
               "output001/output.vtm", "checkpoint000.dat"):
        res[nm.replace("/", "__").replace(".", "_")] = np.frombuffer(open(os.path.join(d, nm), "rb").read(), np.uint8)
    res["stdout"] = np.frombuffer(out.encode(), np.uint8)
    shutil.rmtree(d)
    np.savez_compressed(os.path.join(HERE, "ref_vtk8.npz"), **res)
    print("ref_vtk8 done")


if __name__ == "__main__":
    which = sys.argv[1:] or ["point", "tgv16", "uneven", "restart16", "turb16", "vtk8", "tgv64"]
    for w in which:
        globals()["gen_" + w]()
