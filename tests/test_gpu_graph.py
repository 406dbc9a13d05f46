"""hipGraph replay of the rlm plans (rlm_plan_set_graph), the multi-size plan in one graph
launch (rlm_multi_*), and the tridiagonalisation autotuner. A graph is replayed several
times with the program's own eager GPU work and synchronisations between the replays, and
every replay is checked against the eager result (the same kernels: identical values)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eager_work():
    import torch
    a = torch.randn(512, 512, device="cuda")
    (a @ a).sum().item()
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", [0, 1])
def test_plan_graph_replays_match_eager(precision):
    from rlm_amd import RlmPlan
    print(f"[test_plan_graph_replays_match_eager] precision {precision}", flush=True)
    n, g, S = 64, 8, 4
    eager = RlmPlan(n, 1.0, 16.0, g, 3, samples_per_batch=S, precision=precision)
    ref = eager.run(0, S)
    ref5 = eager.run(5, S)                 # a polynomial jump to the first sample
    eager.close()
    plan = RlmPlan(n, 1.0, 16.0, g, 3, samples_per_batch=S, precision=precision)
    plan.set_graph(True)
    for first, want in ((0, ref), (5, ref5), (0, ref)):
        plan.load(first, S)
        for rep in range(3):
            plan.execute()
            got = plan.fetch()
            for a, b in zip(got, want):
                assert np.array_equal(a, b), (first, rep)
            assert plan.stage_times() == {} and plan.elapsed_ms() > 0.0
            _eager_work()
    plan.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_size_one_launch_matches_plans(precision):
    from rlm_amd import RlmPlan, RlmMulti
    sizes, g, S, seed = [64, 128, 192, 256], 4, 2, 11
    print(f"[test_multi_size_one_launch_matches_plans] {sizes} precision {precision}", flush=True)
    refs = []
    for n in sizes:
        p = RlmPlan(n, 1.0, 16.0, g, seed, samples_per_batch=S, precision=precision)
        refs.append(p.run(0, S))
        p.close()
    m = RlmMulti(sizes, 1.0, 16.0, g, seed, samples_per_batch=S, precision=precision)
    m.load(0, S)
    for rep in range(3):
        m.execute()
        for k in range(len(sizes)):
            for a, b in zip(m.fetch(k), refs[k]):
                assert np.array_equal(a, b), (sizes[k], rep)
        assert m.elapsed_ms() > 0.0
        _eager_work()
    m.close()


def test_autotuner_choice_and_cache(tmp_path, monkeypatch):
    import torch
    from rlm_amd import heevd_batched, tuned_kernel, HIP_DIAG_FP64
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("RLM_TUNE_CACHE", str(cache))
    monkeypatch.delenv("RLM_HETRD", raising=False)
    monkeypatch.delenv("RLM_TUNE", raising=False)
    n = 83                                   # a size no other test solves
    print(f"[test_autotuner_choice_and_cache] n={n}", flush=True)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((8, n, n)) + 1j * rng.standard_normal((8, n, n))
    H = (X + X.conj().transpose(0, 2, 1)) / 2
    w, _ = heevd_batched(torch.from_numpy(H).cuda())
    name = tuned_kernel(n, HIP_DIAG_FP64)
    assert name in ("resident", "lower", "lower8", "fused", "blocked"), name
    lines = cache.read_text().split("\n")
    assert any(l.split()[1:] == [str(n), "0", name] for l in lines if l.strip()), lines
    assert np.abs(w.cpu().numpy() - np.linalg.eigvalsh(H)).max() < 1e-11
