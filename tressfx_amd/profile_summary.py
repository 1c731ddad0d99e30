"""Summaries of rocprofv3 CSV output for one kernel: duration statistics from
a --kernel-trace --stats run and HBM bytes per dispatch from --pmc runs of
FETCH_SIZE and WRITE_SIZE, calibrated against a device copy of known size.

    python -m tressfx_amd.profile_summary stats  KERNEL_STATS_CSV KERNEL_SUBSTRING
    python -m tressfx_amd.profile_summary trace  KERNEL_TRACE_CSV KERNEL_SUBSTRING [LAST_N]
    python -m tressfx_amd.profile_summary traffic FETCH_CSV WRITE_CSV KERNEL_SUBSTRING COPY_BYTES OUT_JSON [LAST_N]
    python -m tressfx_amd.profile_summary digest
    python -m tressfx_amd.profile_summary counters COUNTER_CSV KERNEL_SUBSTRING [LAST_N]

`trace` prints, per kernel whose name holds the substring, the duration
statistics of all its dispatches and of its last LAST_N dispatches (the
launches a bench leg timed, when that leg ran last).

`counters` prints, for the dispatches of the kernel, the mean per dispatch of
every counter of a --pmc run and the ratios read from SQ counters: VALU and
LDS instructions per wave, VALU and LDS issue per wave-cycle, LDS bank
conflicts per LDS issue cycle.

Counter values are KiB per dispatch (with LAST_N, the mean over the kernel's
last LAST_N dispatches: a leg that ran last).  The copy kernel of the same run
(__amd_rocclr_copyBuffer, COPY_BYTES read and written) gives the scale factor
counter_bytes / true_bytes for each counter; a kernel's traffic is its counter
bytes divided by that factor.  A traffic record carries the digest of the
fused kernel's sources it was measured on, so a reader can tell whether it
still describes the kernel that is built.
"""
import csv
import hashlib
import json
import os
import statistics
import sys

FUSED_SOURCES = ("csrc/tfx_fused_kernel.hip", "csrc/tfx_device.h", "build.py")


def fused_source_digest():
    """sha256 over the sources and build flags that define k_fused_substep."""
    h = hashlib.sha256()
    here = os.path.dirname(os.path.abspath(__file__))
    for rel in FUSED_SOURCES:
        with open(os.path.join(here, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def kernel_stats(path, kernel):
    out = []
    for r in _rows(path):
        if kernel in r["Name"]:
            out.append({"name": r["Name"], "calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
                        "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3,
                        "total_ms": float(r["TotalDurationNs"]) / 1e6})
    return out


def trace_stats(path, kernel, last_n=0):
    """Per kernel: duration statistics (us) of every dispatch and of the last
    `last_n` dispatches, in dispatch order."""
    per = {}
    for r in _rows(path):
        if kernel in r["Kernel_Name"]:
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            per.setdefault(r["Kernel_Name"], []).append((int(r["Dispatch_Id"]), d))
    out = []
    for name, ds in per.items():
        ds.sort()
        durs = [d for _, d in ds]

        def st(v):
            return {"calls": len(v), "avg_us": statistics.mean(v), "min_us": min(v), "max_us": max(v)}
        rec = {"name": name, "all": st(durs)}
        if 0 < last_n <= len(durs):
            rec[f"last_{last_n}"] = st(durs[-last_n:])
        out.append(rec)
    return out


def _per_dispatch(path, counter):
    """{kernel name: [counter value per dispatch]} in dispatch order (values
    summed over the counter's instances within a dispatch)."""
    acc = {}
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        key = (r["Kernel_Name"], int(r["Dispatch_Id"]))
        acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    per = {}
    for (name, _), v in sorted(acc.items(), key=lambda kv: kv[0][1]):
        per.setdefault(name, []).append(v)
    return per


def traffic(fetch_csv, write_csv, kernel, copy_bytes, last_n=0):
    res = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        per = _per_dispatch(path, counter)
        copies = [v for n, vs in per.items() if "copyBuffer" in n for v in vs]
        # the calibration copies are the largest copyBuffer dispatches
        copies = sorted(copies)[-3:]
        if not copies:
            raise SystemExit(f"{path}: no copy kernel dispatches to calibrate {counter}")
        factor = statistics.mean(copies) * 1024.0 / copy_bytes
        ks = [v for n, vs in per.items() if kernel in n for v in vs]
        if not ks:
            raise SystemExit(f"{path}: no dispatches of {kernel}")
        if last_n > 0:
            names = {n for n in per if kernel in n}
            if len(names) != 1 or last_n > len(ks):
                raise SystemExit(f"{path}: LAST_N needs one kernel name with at least {last_n} dispatches")
            ks = ks[-last_n:]
        res[counter] = {"kib_per_dispatch_mean": statistics.mean(ks), "dispatches": len(ks),
                        "copy_kib": statistics.mean(copies), "factor": factor,
                        "bytes_per_dispatch": statistics.mean(ks) * 1024.0 / factor}
    res["bytes_per_launch"] = res["FETCH_SIZE"]["bytes_per_dispatch"] + res["WRITE_SIZE"]["bytes_per_dispatch"]
    res["kernel"] = kernel
    res["copy_bytes"] = copy_bytes
    res["source_digest"] = fused_source_digest()
    return res


def counters(path, kernel, last_n=0):
    names = sorted({r["Counter_Name"] for r in _rows(path)})
    out = {}
    for c in names:
        ks = [v for n, vs in _per_dispatch(path, c).items() if kernel in n for v in vs]
        if last_n > 0:
            ks = ks[-last_n:]
        if ks:
            out[c] = {"mean_per_dispatch": statistics.mean(ks), "dispatches": len(ks)}
    m = {c: v["mean_per_dispatch"] for c, v in out.items()}
    ratios = {}
    if m.get("SQ_WAVES"):
        for c, label in (("SQ_INSTS_VALU", "valu_insts_per_wave"), ("SQ_INSTS_LDS", "lds_insts_per_wave")):
            if c in m:
                ratios[label] = m[c] / m["SQ_WAVES"]
    if m.get("SQ_WAVE_CYCLES"):
        for c, label in (("SQ_ACTIVE_INST_VALU", "valu_issue_per_wave_cycle"),
                         ("SQ_ACTIVE_INST_LDS", "lds_issue_per_wave_cycle"),
                         ("SQ_WAIT_INST_LDS", "lds_wait_per_wave_cycle")):
            if c in m:
                ratios[label] = m[c] / m["SQ_WAVE_CYCLES"]
    if m.get("SQ_ACTIVE_INST_LDS") and "SQ_LDS_BANK_CONFLICT" in m:
        ratios["lds_bank_conflict_per_lds_issue_cycle"] = m["SQ_LDS_BANK_CONFLICT"] / m["SQ_ACTIVE_INST_LDS"]
    return {"kernel": kernel, "counters": out, "ratios": ratios}


def main(argv):
    if argv[:1] == ["stats"]:
        for k in kernel_stats(argv[1], argv[2]):
            print(json.dumps(k))
    elif argv[:1] == ["trace"]:
        for k in trace_stats(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 0):
            print(json.dumps(k))
    elif argv[:1] == ["traffic"]:
        r = traffic(argv[1], argv[2], argv[3], int(argv[4]), int(argv[6]) if len(argv) > 6 else 0)
        with open(argv[5], "w") as f:
            json.dump(r, f, indent=1)
        print(json.dumps(r))
    elif argv[:1] == ["counters"]:
        print(json.dumps(counters(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 0), indent=1))
    elif argv[:1] == ["digest"]:
        print(fused_source_digest())
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
