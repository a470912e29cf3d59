"""Per-kernel summary of rocprofv3 counter passes (`--pmc ... --output-format csv`).

Two passes of the same program give, per kernel (dispatches with the same
name and grid averaged):

  pass A  FETCH_SIZE, SQ_WAVES, SQ_WAVE_CYCLES, SQ_BUSY_CYCLES, SQ_ACTIVE_INST_VALU,
          SQ_THREAD_CYCLES_VALU, SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_WAIT_INST_LDS, GRBM_GUI_ACTIVE
  pass B  WRITE_SIZE, TCC_HIT_sum, TCC_MISS_sum, SQ_INSTS_VMEM_RD, SQ_INSTS_VMEM_WR,
          SQ_LDS_BANK_CONFLICT, SQ_ACTIVE_INST_LDS, SQ_INSTS_SALU, SQ_WAIT_INST_ANY,
          SQ_ACTIVE_INST_ANY, SQ_INSTS_SMEM, GRBM_GUI_ACTIVE

Derived figures (rocprofv3's own metric definitions; GRBM_GUI_ACTIVE is summed
over the 8 XCDs, so the kernel's cycle window is that sum / 8; SQ cycle counters
are in quad-cycles):

  waves per SIMD     SQ_WAVE_CYCLES / window / CUs          (OccupancyPercent * 32 / 400 / 4)
  VALU busy          SQ_ACTIVE_INST_VALU / CUs / window       (VALUBusy / 100)
  lane utilization   SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU * 64)   (VALUUtilization / 100)
  L2 hit rate        TCC_HIT / (TCC_HIT + TCC_MISS)
  HBM bytes          FETCH_SIZE * 1024 * fetch_scale, WRITE_SIZE * 1024 * write_scale, the scales
                     being what the counters under-read for the plain copy kernel (copy16) over a
                     known size (bench.py --copy-calibration-mib)

For kernels shorter than about 20 us the window is mostly dispatch overhead,
so their busy and occupancy figures read low.

    python -m lux_amd.counters A.csv B.csv --fetch-scale 2 --write-scale 1 \\
        --rays 4194304 --triangles 1048576 --json profiles/counters.json --traffic profiles/traffic.json
"""

import argparse
import collections
import csv
import json
from pathlib import Path

CUS = 256
XCDS = 8


def _short(name):
    n = name.replace("lux::pt::(anonymous namespace)::", "pt::").replace("lux::(anonymous namespace)::", "")
    return n.replace("void ", "").split("(")[0]


def load(path):
    disp = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        d = disp.setdefault(r["Dispatch_Id"], {
            "name": _short(r["Kernel_Name"]), "grid": int(r["Grid_Size"]), "wg": int(r["Workgroup_Size"]),
            "vgpr": int(r["VGPR_Count"]), "lds": int(r["LDS_Block_Size"]),
            "dur_us": (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3, "c": {}})
        d["c"][r["Counter_Name"]] = float(r["Counter_Value"])
    groups = collections.OrderedDict()
    for d in disp.values():
        groups.setdefault((d["name"], d["grid"]), []).append(d)
    return groups


def _mean(ds, k):
    v = [d["c"][k] for d in ds if k in d["c"]]
    return sum(v) / len(v) if v else float("nan")


def summarize(path_a, path_b, fetch_scale=2.0, write_scale=1.0):
    ga, gb = load(path_a), load(path_b)
    out = collections.OrderedDict()
    for key, ds in ga.items():
        name, grid = key
        db = gb.get(key, [])
        win = _mean(ds, "GRBM_GUI_ACTIVE") / XCDS
        waves = _mean(ds, "SQ_WAVES")
        wc = _mean(ds, "SQ_WAVE_CYCLES")
        act = _mean(ds, "SQ_ACTIVE_INST_VALU")
        e = {
            "kernel": name, "grid": grid, "workgroup": ds[0]["wg"], "vgpr": ds[0]["vgpr"], "lds": ds[0]["lds"],
            "dispatches": len(ds), "kernel_us": round(sum(d["dur_us"] for d in ds) / len(ds), 2),
            "waves": int(waves),
            "waves_per_simd": round(wc / win / CUS, 3),
            "valu_busy_frac": round(act / CUS / win, 4),
            "lane_utilization": round(_mean(ds, "SQ_THREAD_CYCLES_VALU") / (act * 64), 4) if act else None,
            "valu_insts_per_wave": round(_mean(ds, "SQ_INSTS_VALU") / waves, 1),
            "lds_insts_per_wave": round(_mean(ds, "SQ_INSTS_LDS") / waves, 1),
            "fetch_bytes": int(_mean(ds, "FETCH_SIZE") * 1024 * fetch_scale),
        }
        if db:
            h, m = _mean(db, "TCC_HIT_sum"), _mean(db, "TCC_MISS_sum")
            e.update({
                "write_bytes": int(_mean(db, "WRITE_SIZE") * 1024 * write_scale),
                "l2_hit_rate": round(h / (h + m), 4) if h + m else None,
                "vmem_rd_per_wave": round(_mean(db, "SQ_INSTS_VMEM_RD") / waves, 1),
                "vmem_wr_per_wave": round(_mean(db, "SQ_INSTS_VMEM_WR") / waves, 1),
                "salu_per_wave": round(_mean(db, "SQ_INSTS_SALU") / waves, 1),
                "wait_any_frac": round(_mean(db, "SQ_WAIT_INST_ANY") / wc, 4),
                "lds_bank_conflict_frac": round(_mean(db, "SQ_LDS_BANK_CONFLICT") /
                                                (_mean(db, "GRBM_GUI_ACTIVE") / XCDS) / CUS, 5),
            })
        out[f"{name}@{grid}"] = e
    return out


COLUMNS = [("kernel_us", "us"), ("waves_per_simd", "waves/SIMD"), ("valu_busy_frac", "VALU busy"),
           ("lane_utilization", "lanes/instr"), ("valu_insts_per_wave", "VALU/wave"), ("l2_hit_rate", "L2 hit"),
           ("fetch_bytes", "fetch MB"), ("write_bytes", "write MB")]


def table(summary, min_us=5.0):
    rows = ["| kernel @ grid | " + " | ".join(c for _, c in COLUMNS) + " |",
            "|---" * (len(COLUMNS) + 1) + "|"]
    for key, e in sorted(summary.items(), key=lambda kv: -kv[1]["kernel_us"] * kv[1]["dispatches"]):
        if e["kernel_us"] < min_us:
            continue
        cells = []
        for f, _ in COLUMNS:
            v = e.get(f)
            if v is None:
                cells.append("-")
            elif f.endswith("_bytes"):
                cells.append(f"{v / 1e6:.1f}")
            elif isinstance(v, float):
                cells.append(f"{v:.3g}")
            else:
                cells.append(str(v))
        rows.append(f"| {key} | " + " | ".join(cells) + " |")
    return "\n".join(rows)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("pass_a")
    ap.add_argument("pass_b")
    ap.add_argument("--fetch-scale", type=float, default=2.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--rays", type=int, default=0)
    ap.add_argument("--triangles", type=int, default=0)
    ap.add_argument("--json", help="merge the ray-query kernels' entries into this file")
    ap.add_argument("--traffic", help="merge trace_any / trace_closest HBM bytes into this file")
    ap.add_argument("--source", default="")
    a = ap.parse_args()
    s = summarize(a.pass_a, a.pass_b, a.fetch_scale, a.write_scale)
    print(table(s))
    picks = {"k_trace_any": "trace_any", "k_trace_closest<true>": "trace_closest"}
    if a.json:
        p = Path(a.json)
        table_ = json.loads(p.read_text()) if p.exists() else {}
        for key, e in s.items():
            name = e["kernel"]
            if name in picks and e["grid"] == a.rays:
                table_[name.split("<")[0]] = dict(e, rays=a.rays, triangles=a.triangles, source=a.source)
        p.write_text(json.dumps(table_, indent=1) + "\n")
    if a.traffic:
        p = Path(a.traffic)
        table_ = json.loads(p.read_text()) if p.exists() else {}
        for key, e in s.items():
            if e["kernel"] in picks and e["grid"] == a.rays and "write_bytes" in e:
                table_[picks[e["kernel"]]] = {
                    "rays": a.rays, "triangles": a.triangles, "fetch_bytes": e["fetch_bytes"],
                    "write_bytes": e["write_bytes"], "fetch_scale": a.fetch_scale, "write_scale": a.write_scale,
                    "bytes_per_launch": e["fetch_bytes"] + e["write_bytes"], "source": a.source}
        p.write_text(json.dumps(table_, indent=1) + "\n")


if __name__ == "__main__":
    main()
