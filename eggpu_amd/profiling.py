"""Counter calibration for rocprofv3 --pmc runs of bench.py and
eggpu_amd.benchmarks: one plain device copy of known size beside the
workload, and the HBM bytes per launch of named kernels from a FETCH_SIZE
run and a WRITE_SIZE run, each counter scaled by what it read for the copy.
"""
import os


def calib_copy(nbytes):
    """One plain device-to-device copy of `nbytes` through the HIP runtime
    (its blit kernel), so rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
    this command read a kernel of known traffic beside the iteration's."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMalloc failed for the calibration copy")
    try:
        rc = hip.hipMemset(b, 1, ctypes.c_size_t(nbytes)) or \
            hip.hipMemcpy(a, b, ctypes.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize()  # 3: device to device
        if rc != 0:
            raise SystemExit(f"calibration copy failed (hip error {rc})")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def _pmc_rows(path, counter):
    """{kernel name: [(dispatch id, value)]} of one counter from a rocprofv3
    --pmc output directory (*_counter_collection.csv files)."""
    import csv
    import glob

    files = [path] if os.path.isfile(path) else glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                          recursive=True)
    acc = {}
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                key = (r["Kernel_Name"], int(r["Dispatch_Id"]))
                acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for (name, disp), v in sorted(acc.items(), key=lambda kv: kv[0][1]):
        out.setdefault(name, []).append((disp, v))
    return out


def pmc_traffic(fetch_dir, write_dir, copy_bytes, kernels):
    """HBM bytes per launch of the named kernels (summed: the launches of one
    iteration) from a FETCH_SIZE run and a WRITE_SIZE run of this script
    with --calib-copy, each counter scaled by what it read for the copy.
    The copy is the run's first launch of the runtime's copy kernel (the
    library's own copies come later).  Launches that found the stop flag
    already set return at once; only launches with at least half of the
    kernel's largest reading are averaged."""
    fetch = _pmc_rows(fetch_dir, "FETCH_SIZE")
    write = _pmc_rows(write_dir, "WRITE_SIZE")

    def first(d, part):
        for k, v in d.items():
            if part in k and v:
                return v[0][1]
        return None

    def mean_full(d, part):
        for k, v in d.items():
            if part in k and v:
                top = max(x for _, x in v)
                full = [x for _, x in v if x >= 0.5 * top]
                return sum(full) / len(full), len(full), k
        return None, 0, None

    cf, cw = first(fetch, "copyBuffer"), first(write, "copyBuffer")
    if None in (cf, cw):
        raise SystemExit(f"no calibration copy in the counter files (fetch {cf}, write {cw})")
    f_scale, w_scale = copy_bytes / cf, copy_bytes / cw
    per = []
    for part in kernels:
        kf, nf, kname = mean_full(fetch, part)
        kw, nw, _ = mean_full(write, part)
        if None in (kf, kw):
            raise SystemExit(f"no counters for {part}")
        per.append({"kernel": kname, "dispatches": {"fetch": nf, "write": nw}, "FETCH_SIZE_mean": kf,
                    "WRITE_SIZE_mean": kw, "hbm_read_bytes": kf * f_scale, "hbm_write_bytes": kw * w_scale})
    rd = sum(p["hbm_read_bytes"] for p in per)
    wr = sum(p["hbm_write_bytes"] for p in per)
    return {"kernel": " + ".join(p["kernel"] for p in per), "copy_bytes": copy_bytes, "copy_FETCH_SIZE": cf,
            "copy_WRITE_SIZE": cw, "fetch_scale_bytes_per_unit": f_scale, "write_scale_bytes_per_unit": w_scale,
            "hbm_read_bytes_per_launch": rd, "hbm_write_bytes_per_launch": wr, "hbm_bytes_per_launch": rd + wr,
            "per_kernel": per}
