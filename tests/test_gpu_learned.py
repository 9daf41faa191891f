"""The learned preconditioner inside flexible CG, judged by the true FP64
residual (computed here in numpy) and by the held-out acceptance solve that
the library runs after training."""
import numpy as np
import pytest

from hydro_amd import capi, workloads
from hydro_amd.hydro import Hydro, Poisson
from tests.test_gpu_poisson import apply

pytestmark = pytest.mark.gpu

N = 128


@pytest.fixture(scope="module")
def problem():
    kx, ky, rhs = workloads.poisson_problem(N, seed=5)
    return kx, ky, rhs


def true_residual(kx, ky, rhs, p):
    b = rhs - rhs.mean()
    return np.linalg.norm(b - apply(kx, ky, p, N, N)) / np.linalg.norm(b)


def test_learned_multilevel_net_beats_jacobi(gpu, problem):
    """The tied (symmetric positive definite) network alone as the
    preconditioner, M = S N S: fewer CG iterations than Jacobi-PCG, the
    method without it, on the held-out solve the library runs after
    training (both to a 1e-8 true FP64 residual).  It is used by the solvers
    only if it is also faster, which the acceptance flag reports."""
    kx, ky, rhs = problem
    print(f"case: learned standalone {N}x{N}", flush=True)
    ps = Poisson(N, N, kx, ky)
    st = ps.train(steps=150, lr=1e-3, rhs_count=2, mg_iters=6,
                  flags=capi.HG_TRAIN_TIED | capi.HG_TRAIN_STANDALONE)
    print(f"train: pairs {st.pairs} loss {st.loss_first:.3f} -> {st.loss_last:.3f}; held-out "
          f"{st.val_iters} it {st.val_ms:.2f} ms vs jacobi-pcg {st.val_iters_baseline} it "
          f"{st.val_ms_baseline:.2f} ms; accepted {st.accepted}", flush=True)
    assert 0 < st.val_iters < st.val_iters_baseline
    assert bool(st.accepted) == (st.val_ms < st.val_ms_baseline)
    p, s1 = ps.solve(rhs, solver="fcg_learned", tol=1e-8, precision=1)
    assert s1.converged and true_residual(kx, ky, rhs, p) <= 1.05e-8
    ps.close()


def test_learned_correction_is_gated(gpu, problem):
    """V-cycle plus learned correction: used only when the held-out solve
    shows fewer iterations than the V-cycle alone; the solve is correct
    either way."""
    kx, ky, rhs = problem
    print(f"case: learned additive {N}x{N}", flush=True)
    ps = Poisson(N, N, kx, ky)
    st = ps.train(steps=100, lr=1e-3, rhs_count=2, mg_iters=6, flags=capi.HG_TRAIN_TIED)
    print(f"train: accepted {st.accepted}; held-out {st.val_iters} it vs fcg+V-cycle "
          f"{st.val_iters_baseline} it", flush=True)
    assert bool(st.accepted) == (st.val_iters < st.val_iters_baseline and
                                 st.val_ms < st.val_ms_baseline)
    p, s1 = ps.solve(rhs, solver="fcg_learned", tol=1e-8, precision=1)
    _, s_mg = ps.solve(rhs, solver="fcg_mg", tol=1e-8, precision=1)
    assert s1.converged and true_residual(kx, ky, rhs, p) <= 1.05e-8
    if not st.accepted:
        assert s1.iterations == s_mg.iterations
    ps.close()


def test_train_pressure_net_in_simulation(gpu):
    """Training on the simulation's own pressure operator, then stepping on."""
    print("case: rt 64x128 fcg_learned training", flush=True)
    with Hydro(workloads.rt(64, 128, solver_p="fcg_learned", precision="mixed")) as h:
        h.step()
        st = h.train_pressure_net(steps=60, rhs_count=1, mg_iters=4)
        assert st.pairs == 4 and st.steps == 60
        s = h.step()
        assert s.simple_iterations >= 1
        assert np.all(np.isfinite(h.get(capi.HG_FIELD_PRESSURE)))
    with Hydro(workloads.rt(64, 128, solver_p="cg", precision="fp64")) as h:
        with pytest.raises(Exception):
            h.train_pressure_net(steps=1)
