"""Per-kernel hardware-counter summary of one benchmark run, for profiles/counters.json.

The input is a set of rocprofv3 output directories of the same command, one per counter pass
(`<prefix>_trace` from --kernel-trace, and `<prefix>_fetch`, `_write`, `_tcp`, `_tcc`, `_lds`
from --pmc passes; the counter sets are listed in profiles/sigma_c2_counters.md).  For each
phase of the sigma build (mixed kernel, alpha and beta row passes, transposes) it reports per
build: time, HBM-side bytes (FETCH_SIZE and WRITE_SIZE, calibrated on the bandwidth probe's
copy kernel of known size), L2 -> L1 bytes (TCP_TCC_READ_REQ x the request size the copy kernel
shows), L2 hit rate, TA busy fraction, L1 pending-stall fraction, instruction counts and the
clock.  A phase with several launches per build (the row passes' ket blocks) is summed over
them.

    python -m sbd_amd.pmc_summary <prefix> c2/fp64 [profiles/counters.json]

where <prefix>_trace, <prefix>_fetch, ... are the rocprofv3 output folders of one workload and
precision (their -d arguments).
"""
import collections
import csv
import json
import os
import sys

PHASES = {"k_sigma_mixed": "mixed", "k_sigma_rows<double, true": "alpha_rows",
          "k_sigma_rows<double, false": "beta_rows", "k_sigma_rows<float, true": "alpha_rows",
          "k_sigma_rows<float, false": "beta_rows", "k_transpose<": "transpose",
          "k_add_transposed": "add_transposed", "k_probe_copy": "probe_copy", "k_probe_l2": "probe_l2"}
PROBE_BYTES = 4 * 2**30  # the copy kernel reads and writes 4 GiB per launch
SOURCE = ("profiles/sigma_c2_counters.md: rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, TCP/TA, "
          "TCC/SQ, LDS) over bench.py --steps 2 --warmup 0 --no-solve --no-cpu-baseline, calibrated "
          "on the bandwidth probe's copy kernel")


def phase(name):
    for k, v in PHASES.items():
        if k in name:
            return v
    return None


def summarize(prefix):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in ("fetch", "write", "tcp", "tcc", "lds"):
        f = f"{prefix}_{d}/pmc_counter_collection.csv"
        if not os.path.exists(f):
            continue
        for r in csv.DictReader(open(f)):
            k = phase(r["Kernel_Name"])
            if k:
                agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(f"{prefix}_trace/pmc_kernel_trace.csv")):
        k = phase(r["Kernel_Name"])
        if k:
            dur[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    builds = len(dur["mixed"])
    launches = {k: len(v) / builds for k, v in dur.items()}
    # (mean per dispatch) x (dispatches per build); a counter taken in several passes (GRBM)
    # appears once per pass and dispatch, so the mean is the right per-dispatch figure
    m = {k: {c: sum(v) / len(v) * (1 if k.startswith("probe") else launches[k]) for c, v in cs.items()}
         for k, cs in agg.items()}
    cp = m["probe_copy"]
    f_fetch = PROBE_BYTES / (cp["FETCH_SIZE"] * 1024)
    f_write = PROBE_BYTES / (cp["WRITE_SIZE"] * 1024)
    req_bytes = PROBE_BYTES / cp["TCP_TCC_READ_REQ_sum"]
    out = {}
    for k in ("mixed", "alpha_rows", "beta_rows", "transpose", "add_transposed"):
        if k not in m:
            continue
        c = m[k]
        t = sum(dur[k]) / builds
        xcd_cycles = c["GRBM_GUI_ACTIVE"] / 8
        rec = dict(mean_ms=round(t * 1e3, 3),
                   hbm_bytes=int(c["FETCH_SIZE"] * 1024 * f_fetch + c["WRITE_SIZE"] * 1024 * f_write),
                   l2_to_l1_bytes=int(c["TCP_TCC_READ_REQ_sum"] * req_bytes),
                   l2_hit=round(c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4),
                   ta_busy=round(c["TA_BUSY_avr"] / xcd_cycles, 4),
                   tcp_pending_stall=round(c["TCP_PENDING_STALL_CYCLES_sum"] / 256 / xcd_cycles, 4),
                   vmem_rd_instr=int(c["SQ_INSTS_VMEM_RD"]), lds_instr=int(c["SQ_INSTS_LDS"]),
                   valu_instr=int(c["SQ_INSTS_VALU"]), clock_ghz=round(xcd_cycles / t / 1e9, 3),
                   launches_per_build=launches[k], source=SOURCE)
        if "SQ_LDS_BANK_CONFLICT" in c:
            rec["lds_bank_conflict_frac"] = round(c["SQ_LDS_BANK_CONFLICT"] / 256 / xcd_cycles, 4)
            rec["lds_active_frac"] = round(c.get("SQ_LDS_IDX_ACTIVE", 0) / 256 / xcd_cycles, 4)
        out[k] = rec
    calib = dict(fetch_factor=round(f_fetch, 4), write_factor=round(f_write, 4),
                 tcp_tcc_read_req_bytes=round(req_bytes, 2),
                 probe_copy_ms=round(sum(dur["probe_copy"]) / len(dur["probe_copy"]) * 1e3, 4))
    return out, calib


def main(argv):
    prefix, key = argv[1], argv[2]
    path = argv[3] if len(argv) > 3 else os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                      "profiles", "counters.json")
    out, calib = summarize(prefix)
    allc = json.load(open(path)) if os.path.exists(path) else {}
    allc[key] = out
    allc.setdefault("_calibration", {})[key] = calib
    json.dump(allc, open(path, "w"), indent=1)
    print(json.dumps({"phases": out, "calibration": calib}, indent=1))


if __name__ == "__main__":
    main(sys.argv)
