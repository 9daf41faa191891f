"""Short, fixed workloads to run under rocprofv3 (kernel traces and counters).

    python -m hydro_amd.profile_targets step      # RT 1024x2048 headline steps
    python -m hydro_amd.profile_targets kernels   # a 256 MiB device copy (counter
                                                  # calibration), then the step's
                                                  # dominant kernels
    python -m hydro_amd.profile_targets net       # the BF16 MFMA network alone
    python -m hydro_amd.profile_targets step3d    # 3D cavity 128^3 steps (hg3_*)

Each prints what it ran; none writes files.
"""
import ctypes
import sys
import time


def _rt(solver_p="fcg_mg"):
    from hydro_amd import workloads
    from hydro_amd.hydro import Hydro
    h = Hydro(workloads.rt(1024, 2048, solver_p=solver_p, solver_v="lu_rb", precision="mixed"))
    workloads.set_rt_state(h)
    return h


def step(n=20, warmup=3):
    """The benchmark's headline steps: `warmup` steps, then `n` steps from
    the same Rayleigh-Taylor state.  Each step runs KForceFaces once, so the
    timed steps are the dispatches from KForceFaces number warmup + 1 on."""
    h = _rt()
    for _ in range(warmup):
        h.step()
    h.timing()
    t0 = time.perf_counter()
    its = []
    for i in range(n):
        st = h.step()
        its.append(st.pressure_solve_iters_total)
    h.timing()
    print(f"{warmup} warm-up steps, then {n} steps: {1e3 * (time.perf_counter() - t0) / n:.3f} "
          f"ms/step, pressure iterations per step {sum(its) / n:.2f} {its}", flush=True)
    h.close()


def kernels(reps=5):
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):  # device to device
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("calibration copy: 3 x 256 MiB", flush=True)
    from hydro_amd.hydro import Poisson
    h = _rt()
    h.step()
    for k in ("KAdvStep", "KMomentum"):
        print(f"{k}: {h.time_kernel(k, reps=reps) * 1e3:.2f} us", flush=True)
    kf, rhs = h.get(16), h.get(17)
    h.close()
    nx, ny = 1024, 2048
    ps = Poisson(nx, ny, kf[:(nx + 1) * ny], kf[(nx + 1) * ny:])
    ps.solve(rhs, solver="fcg_mg", tol=1e-8, precision=1)
    for k in ("KUp<double, true>", "KDown<double>", "KCoarseCycle"):
        print(f"{k}: {ps.time_kernel(k, reps=reps) * 1e3:.2f} us", flush=True)
    ps.close()


def net(reps=5):
    from hydro_amd.hydro import Net
    n = Net(1024, 2048)
    print(n.time(reps=reps), flush=True)
    n.close()


def step3d(n=128, steps=5, warmup=2, solver_p="fcg_mg"):
    """The 3D step: cavity n^3, FP64, jacobi velocity; pressure by solver_p
    (fcg_mg: flexible CG with the FP32 V-cycle; cg: Jacobi-PCG), to 1e-8."""
    from hydro_amd import workloads
    from hydro_amd.hydro import Hydro3
    h = Hydro3(workloads.cavity3d(n, n, n, solver_p=solver_p, solver_v="jacobi"))
    for _ in range(warmup):
        h.step()
    t0 = time.perf_counter()
    st = [h.step() for _ in range(steps)]
    print(f"3D cavity {n}^3: {warmup} warm-up steps, then {steps} steps: "
          f"{1e3 * (time.perf_counter() - t0) / steps:.2f} ms/step, SIMPLE "
          f"{[s.simple_iterations for s in st]}, pressure iterations "
          f"{[s.pressure_solve_iters_total for s in st]}", flush=True)
    h.close()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "step"
    targets = {"step": step, "kernels": kernels, "net": net, "step3d": step3d,
               "step3d_cg": lambda: step3d(solver_p="cg")}
    targets[what]()
