"""Cycles per chain stage on the GPU, from the profiling build.

    make -C e1039-gpu_amd/csrc timing
    python e1039-gpu_amd/stage_cycles.py [n_events] [mean_hits | occ8] [pruned] [high]

Runs one warm-up and one measured launch of the chain on seeded events
(hits/event ~ Poisson(mean_hits) capped at 1000, or 8 % noise occupancy per
channel with occ8) through libe1039or_timing.so
and prints each stage's share of the summed wave cycles.  With "pruned" the
shipped YZ unit pruner is on; with "high" the caps are orchain.HIGH_OCCUPANCY_CAPS
and the hit count is capped at 4000 instead of 1000.
"""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

STAGES = ["load+reduce", "X pairs", "XZ", "YZ", "cleaning", "global", "outputs"]
SUBS = ["YZ window lists", "YZ pass 1 (first fit)", "YZ pass 2 (seeds)", "YZ pass 3 (tails)", "YZ winners",
        "YZ unit path", "st1 window lists", "st1 pass A (y fits)", "st1 pass B + write",
        "YZ unit scoring (before pass 1)", "YZ candidates (count)", "YZ units scored (count)",
        "YZ units kept (count)", "YZ candidates of kept units (count)", "YZ first-cut survivors (count)",
        "YZ pass 3 sign-free tests (in pass 3)", "YZ pass 3 survivors of them (count)"]


def main():
    import numpy as np
    import torch

    import orchain as oc

    n = int(sys.argv[1]) if len(sys.argv) > 1 else 200_000
    wl = sys.argv[2] if len(sys.argv) > 2 else "300"
    # "occ8": 8 % noise occupancy per channel; a number: Poisson mean hits capped at 1000
    high = "high" in sys.argv[3:]
    gen = dict(mult_mode=0, noise_occupancy=0.08) if wl == "occ8" else dict(mult_mode=1, mult_mean=float(wl),
                                                                              mult_cap=4000 if high else 1000)
    lib = oc.load_library(os.path.join(HERE, "libe1039or_timing.so"))
    lib.or_debug_stage_cycles.argtypes = [C.c_void_p, C.c_int]
    lib.or_debug_stage_cycles.restype = C.c_int32
    geo = oc.Geometry.load()
    ev = oc.generate(geo, n, seed=7, **gen)
    params = oc.high_occupancy_params() if high else oc.default_params(max_hits_per_event=1024)
    sess = oc.Session(geo, params, device=0)
    if "pruned" in sys.argv[3:]:
        import pruner
        flat, thr, meta = pruner.load()
        sess.set_pruner_margin(meta.get("margin", float("inf")))
        sess.set_pruner(flat, thr)
        print(f"YZ unit pruner on (keep margin {meta.get('margin', float('inf')):.4f})")
    din = sess.upload(ev)
    dout = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
    sess.run_device(din, dout)
    torch.cuda.synchronize()
    lib.or_debug_sub_cycles.argtypes = [C.c_void_p, C.c_int]
    lib.or_debug_sub_cycles.restype = C.c_int32
    buf = (C.c_ulonglong * 35)()
    sub = (C.c_ulonglong * 18)()
    assert lib.or_debug_stage_cycles(buf, 1) == 0, "not the profiling build"
    lib.or_debug_sub_cycles(sub, 1)
    sess.run_device(din, dout)
    torch.cuda.synchronize()
    desc = "8 % noise occupancy" if wl == "occ8" else f"Poisson({float(wl):g}) hits/event"
    print(f"{n} events, {desc} ({ev.n_hits / n:.0f} hits/event): chain kernels {sess.last_kernel_ms():.1f} ms")
    lib.or_debug_stage_cycles(buf, 0)
    lib.or_debug_sub_cycles(sub, 0)
    a = np.array(list(buf), dtype=np.float64).reshape(5, 7)
    for c in range(5):
        tot = a[c].sum()
        if tot == 0:
            continue
        print(f"tile class {512 << c} hits: {tot:.3e} cycles")
        for k, name in enumerate(STAGES):
            print(f"  {name:12s} {a[c, k] / tot * 100:5.1f} %  {a[c, k] / n:12.0f} cycles/event")
    print("sub-stages (all classes), cycles/event (counts per event for the count rows):")
    for k, name in enumerate(SUBS):
        print(f"  {name:22s} {sub[k] / n:12.0f}")


if __name__ == "__main__":
    main()
