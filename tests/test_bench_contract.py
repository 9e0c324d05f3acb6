"""Host-side pieces of bench.py: the committed traffic summaries are found by
the plan they were measured at, and only by that plan."""
import glob
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_every_traffic_summary_names_its_plan():
    files = sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md")))
    assert files
    for path in files:
        keys = [line.split(":", 1)[0] for line in open(path)]
        assert keys.count("traffic_plan") == 1, path
        assert keys.count("traffic_bytes_per_launch") == 1, path


def _summaries():
    """(kernel, workload, plan line, bytes, file name) of every committed
    traffic summary."""
    out = []
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        name = os.path.basename(path)
        parts = name.split("_")
        vals = dict(line.rstrip("\n").split(": ", 1) for line in open(path)
                    if line.startswith(("traffic_plan:", "traffic_bytes_per_launch:")))
        out.append((parts[0] + "_" + parts[1], parts[2], vals["traffic_plan"],
                    float(vals["traffic_bytes_per_launch"]), name))
    return out


def test_traffic_lookup_matches_the_plan_exactly():
    b = _bench()
    S = _summaries()
    plans = {(k, w, p) for k, w, p, _, _ in S}
    for kernel, wl, plan, nbytes, name in S:
        f = dict(x.split("=") for x in plan.split())
        P = dict(degree=int(f["degree"]), theta=float(f["theta"]), tree_levels=int(f["levels"]),
                 many_count=int(f["many_count"]), node_rule=int(f["nodes"]), theta_decay=float(f["decay"]))
        tile = int(f["tile"])
        assert b.plan_key(P, tile) == plan
        assert b.traffic_from_profiles(kernel, wl, plan) == (nbytes, name)
        # another tile, depth or decay is another plan: no traffic unless a
        # summary was measured at that plan too
        for other in (b.plan_key(P, 6 - tile), b.plan_key(dict(P, tree_levels=P["tree_levels"] + 1), tile),
                      b.plan_key(dict(P, theta_decay=P["theta_decay"] * 0.5), tile)):
            if (kernel, wl, other) not in plans:
                assert b.traffic_from_profiles(kernel, wl, other) == (None, None)
    assert b.traffic_note(None, "x").startswith("null")


def test_more_than_one_gpu_is_refused_with_a_message():
    import subprocess
    import sys

    for argv, env in ((["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "4"})):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, capture_output=True, text=True,
                           env=dict(os.environ, **env), timeout=60)
        assert r.returncode == 2
        assert "one GPU only" in r.stderr
        assert r.stdout == ""
