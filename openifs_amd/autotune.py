"""Autotuner of the blocked CLOUDSC launch, per arithmetic variant and level count.

    python -m openifs_amd.autotune --variant sp mixed --klev 137 91 --columns 2621440

For each (variant, KLEV) it times, on seeded synthetic atmospheres of the requested size:
  * NPROMA (the block length of the IFS fields) x work-group size (64, 128, 256, 512 columns);
  * for the two best (NPROMA, work-group size) pairs, the register dimension -- the minimum waves per SIMD the
    kernel's register allocation must allow (the compiler's choice, or the variant's instantiated bounds: fewer
    registers per lane for more waves in flight, at the price of spills) -- and the kernel with the level count
    compiled in (KLEV 137 and 91).
It picks the fastest launch, and records it in openifs_amd/tuning/cloudsc_tuned.json under the device's
architecture name.  `apply()` (called by openifs_amd.cloudsc when the library is loaded) installs the recorded
launches through cloudsc_tune_set_launch; the recorded NPROMA is advice for hosts that choose their block length
(the IFS sets NPROMA in its namelist).
"""
import argparse
import ctypes
import json
import os
import time

from . import cloudsc as C

VARIANTS = {"dp": 0, "sp": 1, "mixed": 2}
BLOCKS = (64, 128, 256, 512)
NPROMAS = (64, 128, 256, 512)
# waves-per-SIMD bounds instantiated per variant besides 0 (include/cloudsc_tune.h)
WAVES = {"dp": (0, 2), "sp": (0, 4), "mixed": (0, 3, 4)}
KLEV_COMPILED = (137, 91)
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "cloudsc_tuned.json")


def _arch():
    import torch

    return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]


def set_launch(variant, klev, block, waves=0, kspec=0):
    C._check("cloudsc_tune_set_launch",
             C.lib().cloudsc_tune_set_launch(VARIANTS[variant], klev, block, waves, kspec))


def get_launch(variant, klev):
    b, w, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    C._check("cloudsc_tune_get_launch", C.lib().cloudsc_tune_get_launch(VARIANTS[variant], klev, ctypes.byref(b),
                                                                         ctypes.byref(w), ctypes.byref(k)))
    return b.value, w.value, k.value


def set_block(variant, klev, block):
    C._check("cloudsc_tune_set_block", C.lib().cloudsc_tune_set_block(VARIANTS[variant], klev, block))


def get_block(variant, klev):
    return int(C.lib().cloudsc_tune_get_block(VARIANTS[variant], klev))


def time_launches(variant, klev, ncols, nproma, launches, reps=5, seed=4242, params=None):
    """Median kernel time (ms) of each launch (block, waves, kspec) for one NPROMA."""
    import numpy as np
    import torch

    params = params or C.default_params(900.0, klev)
    f = C.BlockedFields(ncols, nproma, klev, variant, "cuda")
    C.synth_device(f, params, seed)
    plude0 = f["plude"].clone()
    stream = torch.cuda.current_stream()
    out = {}
    old = get_launch(variant, klev)
    try:
        for launch in launches:
            set_launch(variant, klev, *launch)
            ts = []
            for i in range(reps + 1):
                f["plude"].copy_(plude0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)
                e1.record(stream)
                torch.cuda.synchronize()
                if i:
                    ts.append(e0.elapsed_time(e1))
            out[tuple(launch)] = float(np.median(ts))
    finally:
        set_launch(variant, klev, *old)
    del f, plude0
    torch.cuda.empty_cache()
    return out


def time_config(variant, klev, ncols, nproma, blocks, reps=5, seed=4242, params=None):
    """Median kernel time (ms) of each work-group size for one NPROMA (compiler's register allocation)."""
    t = time_launches(variant, klev, ncols, nproma, [(b, 0, 0) for b in blocks], reps, seed, params)
    return {b: ms for (b, _, _), ms in t.items()}


def tune(variant, klev, ncols, npromas=NPROMAS, blocks=BLOCKS, reps=5, log=print):
    """Times (NPROMA, block) pairs, then the register bound and compiled-in KLEV on the two best pairs; returns
    the record for the cache."""
    table = []
    for nproma in npromas:
        t = time_config(variant, klev, ncols, nproma, blocks, reps)
        for b, ms in t.items():
            table.append({"nproma": nproma, "block": b, "waves": 0, "klev_compiled": 0, "ms": ms})
            if log:
                log(f"{variant} KLEV={klev} NPROMA={nproma:4d} block={b:4d}: {ms:.3f} ms")
    top = sorted((r for r in table if r["block"] in (128, 256)), key=lambda r: r["ms"])[:2]
    kspecs = (0, 1) if klev in KLEV_COMPILED else (0,)
    for r in top:
        launches = [(r["block"], w, k) for w in WAVES[variant] for k in kspecs if (w, k) != (0, 0)]
        t = time_launches(variant, klev, ncols, r["nproma"], launches, reps)
        for (b, w, k), ms in t.items():
            table.append({"nproma": r["nproma"], "block": b, "waves": w, "klev_compiled": k, "ms": ms})
            if log:
                log(f"{variant} KLEV={klev} NPROMA={r['nproma']:4d} block={b:4d} waves/SIMD={w} "
                    f"klev_compiled={k}: {ms:.3f} ms")
    best = min(table, key=lambda r: r["ms"])
    word = 8 if variant == "dp" else 4
    return {"block": best["block"], "nproma": best["nproma"], "waves": best["waves"],
            "klev_compiled": best["klev_compiled"], "ms": best["ms"], "columns": ncols,
            "gbs": word * (49 * klev + 20) * ncols / best["ms"] / 1e6, "table": table,
            "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S")}


def load(path=CACHE):
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def save(cache, path=CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(cache=None, arch=None):
    """Installs the recorded launches of this device's architecture; returns what was applied."""
    cache = load() if cache is None else cache
    if arch is None:
        try:
            arch = _arch()
        except Exception:
            return {}
    applied = {}
    for variant, per_klev in cache.get(arch, {}).items():
        for klev, rec in per_klev.items():
            launch = (int(rec["block"]), int(rec.get("waves", 0)), int(rec.get("klev_compiled", 0)))
            set_launch(variant, int(klev), *launch)
            applied[(variant, int(klev))] = launch
    return applied


def recommended_nproma(variant, klev, default=128, cache=None, arch="gfx950"):
    cache = load() if cache is None else cache
    rec = cache.get(arch, {}).get(variant, {}).get(str(klev))
    return int(rec["nproma"]) if rec else default


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--variant", nargs="+", default=["mixed"], choices=sorted(VARIANTS))
    ap.add_argument("--klev", type=int, nargs="+", default=[137, 91])
    ap.add_argument("--columns", type=int, default=2621440)
    ap.add_argument("--npromas", type=int, nargs="+", default=list(NPROMAS))
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--cache", default=CACHE)
    args = ap.parse_args()
    cache = load(args.cache)
    arch = _arch()
    for v in args.variant:
        for k in args.klev:
            rec = tune(v, k, args.columns, npromas=tuple(args.npromas), reps=args.reps,
                       log=lambda s: print(s, flush=True))
            cache.setdefault(arch, {}).setdefault(v, {})[str(k)] = rec
            print(f"best {v} KLEV={k}: NPROMA={rec['nproma']} block={rec['block']} waves/SIMD={rec['waves']} "
                  f"klev_compiled={rec['klev_compiled']} {rec['ms']:.3f} ms ({rec['gbs']:.0f} GB/s)", flush=True)
            save(cache, args.cache)


if __name__ == "__main__":
    main()
