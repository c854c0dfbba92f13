"""Speed of the oracle (the repository's C restatement, bit-exact with the
reference) against the reference binary compiled from its own sources
(oracle/_ref, `make -C oracle ref`), on this host's cores: TGV 64^3, 10
steps, the reference at -np 1 (dims 1 1 1) and -np 8 (2 2 2), the oracle on
1 and 8 OpenMP threads.  The reference's time is its own printed loop wall
time (src/cbns3d.cpp:272-286); the oracle's the same loop timed here.
bench.py's cpu_baseline times the oracle on the GPU host, where the
reference does not travel; this ratio says how far that stands in for it.

    python oracle/calibrate_cpu.py        (needs oracle/_ref/cbns3d_ref)
"""
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle import oracle as O  # noqa: E402

MPIEXEC = "/opt/conda/bin/mpiexec"
N, STEPS = 64, 10


def input_txt(n, steps):
    # the 17 positional values of src/read_input.cpp:24-72 (Taylor-Green)
    vals = ["checkpoint", "0.5", "0.0", "20.0", "0.71", "1.4", "0.1", "1600.0", "255.556", "1.0",
            str(n), str(n), str(n), str(steps), str(steps + 1), "1", "1"]
    return "\n".join(vals) + "\n"


def reference(dims):
    d = tempfile.mkdtemp(prefix="calib_")
    with open(os.path.join(d, "input.txt"), "w") as f:
        f.write(input_txt(N, STEPS))
    nproc = dims[0] * dims[1] * dims[2]
    pre = ["setarch", "x86_64", "-R"] if nproc > 1 else []  # see tests/golden/make_golden.py
    out = subprocess.run(pre + [MPIEXEC, "-np", str(nproc), O.REF_BIN_PATH] + [str(x) for x in dims], cwd=d,
                         check=True, capture_output=True, text=True).stdout
    return float(re.search(r"Wall time:\s*([0-9.eE+-]+)", out).group(1))


def oracle(threads):
    O.set_threads(threads)
    b = O.Block(O.make_params(N))
    b.init_tgv()
    t0 = time.perf_counter()
    b.run(STEPS)
    return time.perf_counter() - t0


def main():
    if not os.path.exists(O.REF_BIN_PATH):
        sys.exit("reference binary not built (make -C oracle ref)")
    rate = lambda s: N ** 3 * STEPS / s / 1e6
    res = {}
    for name, f in (("reference -np 1", lambda: reference((1, 1, 1))), ("reference -np 8", lambda: reference((2, 2, 2))),
                    ("oracle 1 thread", lambda: oracle(1)), ("oracle 8 threads", lambda: oracle(8))):
        f()  # warm-up run
        res[name] = min(f() for _ in range(2))
        print(f"{name}: {res[name]:.3f} s = {rate(res[name]):.3f} Mcell-steps/s", flush=True)
    print(f"oracle / reference speed: 1 core {res['reference -np 1'] / res['oracle 1 thread']:.2f}x, "
          f"8 cores {res['reference -np 8'] / res['oracle 8 threads']:.2f}x")


if __name__ == "__main__":
    main()
