"""Writes the autotuner's picks (tuned.json, "picks") as a C header the
library is built with, so its default launch shapes are the tuned ones.

    python gen_tuned.py tuned.json out/tuned_shapes.h

Each workload (single search, batched mix, serving stream) gets keys per
inversion, waves per CU, limb layout and walk form, and for the batched ones
the step budget and the expected keys per virtual lane. A table without a
"picks" section (an older sweep) takes the single-search pick from "pick" and
the measured defaults for the others.
"""

import json
import sys

WORKLOADS = ("single", "batched", "serving")
# what the library used before any mix sweep: the single-search sweep's pick
# everywhere, 200 ms steps and 2^14 keys per lane for a batch, 25 ms and 2^10
# for serving
DEFAULTS = {
    "single": {"batch": 256, "waves_per_cu": 16, "layout_id": 0, "walk_id": 0, "step_ms": 0, "lane_keys_log2": 0},
    "batched": {"batch": 256, "waves_per_cu": 16, "layout_id": 0, "walk_id": 0, "step_ms": 200,
                "lane_keys_log2": 14},
    "serving": {"batch": 256, "waves_per_cu": 16, "layout_id": 0, "walk_id": 0, "step_ms": 25,
                "lane_keys_log2": 10},
}
FIELDS = ("batch", "waves_per_cu", "layout_id", "walk_id", "step_ms", "lane_keys_log2")


def picks(table):
    out = {w: dict(DEFAULTS[w]) for w in WORKLOADS}
    if table and "pick" in table:
        out["single"].update({k: table["pick"][k] for k in FIELDS if k in table["pick"]})
        out["single"]["keys_per_s"] = table["pick"].get("keys_per_s", 0.0)
    for w in WORKLOADS:
        p = (table or {}).get("picks", {}).get(w)
        if p:
            out[w].update({k: p[k] for k in FIELDS if k in p})
            out[w]["keys_per_s"] = p.get("keys_per_s", 0.0)
    return out


def header(pk):
    rows = []
    for i, w in enumerate(WORKLOADS):
        p = pk[w]
        rows.append("    /* %s */ {%d, %d, %d, %d, %d, %d, %.6e}," % (
            w, p["batch"], p["waves_per_cu"], p["layout_id"], p["walk_id"], p["step_ms"], p["lane_keys_log2"],
            float(p.get("keys_per_s") or 0.0)))
    return ("// Generated by gen_tuned.py from tuned.json: the autotuner's launch shapes,\n"
            "// indexed by LXV_TUNED_SINGLE, LXV_TUNED_BATCHED, LXV_TUNED_SERVING.\n"
            "#pragma once\n#include \"lxmf_vanity_hip.h\"\n"
            "static const lxv_tuned_shape kTunedShapes[3] = {\n" + "\n".join(rows) + "\n};\n")


def main():
    src, dst = sys.argv[1], sys.argv[2]
    try:
        with open(src) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = None
    text = header(picks(table))
    with open(dst, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
