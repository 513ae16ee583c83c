"""Text summaries of rocprofv3 result databases (the SQLite files `rocprofv3 -d DIR -o run`
writes), used for the kernel timings and hardware counters kept under profiles/.

    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py kernels RUN_results.db [NAME_PART ...]
    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py counters RUN_results.db [NAME_PART ...]

`kernels`: per kernel name, the number of dispatches, mean / min / max duration (us) and the
share of the summed kernel time (from the `kernels` view: end - start in ns).
`counters`: per kernel name and counter, the number of dispatches and the mean value per
dispatch (from the `counters_collection` view of a `--pmc` run); with `--each` as the first
NAME_PART, every dispatch's value instead, in dispatch order.
NAME_PART filters the kernel names (substring match, any of them).

    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py timeline RUN_results.db [LAST]

`timeline`: per transform of a `--kernel-trace` run (from its first kernel, `hash_clear_kernel`
or, with the column front end, `fold_slab_kernel`, to the next `estimate_kernel`, its last),
the span, the union of kernel busy time and the idle rest; then the last LAST transforms of
each precision (default 1) kernel by kernel: start, duration, grid, workgroup, LDS bytes and
VGPRs per dispatch.

    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py join KERNELS.txt COUNTERS.txt [...]

`join`: one row per kernel from the text outputs above (a `kernels` summary and any number of
`counters` summaries): mean duration, HBM bytes read (2 x FETCH_SIZE KiB: FETCH_SIZE counts 64
bytes per 128-byte read request on MI355X, measured with a copy of known size) and written
(WRITE_SIZE KiB), their rate and share of the 8 TB/s peak, LDS bank-conflict time per CU
(SQ_LDS_BANK_CONFLICT cycles / 256 CUs / 2.4 GHz) and the waves' waiting share
(SQ_WAIT_ANY / SQ_WAVE_CYCLES).  Standard library only.
"""
import collections
import sqlite3
import sys


def _view(cur, candidates):
    names = {r[0] for r in cur.execute(
        "select name from sqlite_master where type in ('table', 'view')")}
    for v in candidates:
        if v in names:
            return v
    raise SystemExit(f"none of {candidates} in the database (has: {sorted(names)[:20]})")


def _columns(cur, view):
    return {r[1]: i for i, r in enumerate(cur.execute(f"pragma table_info({view})"))}


def _short(name):
    return name.split("(")[0][:96]


def _keep(name, parts):
    return not parts or any(p in name for p in parts)


def kernels(path, parts):
    cur = sqlite3.connect(path).cursor()
    view = _view(cur, ("kernels",))
    ix = _columns(cur, view)
    nk = "name" if "name" in ix else "kernel_name"
    by = collections.defaultdict(list)
    for r in cur.execute(f"select * from {view}"):
        by[r[ix[nk]]].append((r[ix["end"]] - r[ix["start"]]) / 1000.0)
    total = sum(sum(v) for v in by.values()) or 1.0
    print(f"{'calls':>5} {'mean_us':>9} {'min_us':>9} {'max_us':>9} {'share':>6}  kernel")
    for name, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        if _keep(name, parts):
            print(f"{len(v):5d} {sum(v) / len(v):9.1f} {min(v):9.1f} {max(v):9.1f} "
                  f"{100.0 * sum(v) / total:5.1f}%  {_short(name)}")


def counters(path, parts):
    cur = sqlite3.connect(path).cursor()
    view = _view(cur, ("counters_collection", "pmc_events", "counters"))
    ix = _columns(cur, view)
    kn = next(c for c in ("kernel_name", "name") if c in ix)
    cn = next(c for c in ("counter_name", "name_counter", "counter") if c in ix)
    vn = next(c for c in ("value", "counter_value") if c in ix)
    if parts and parts[0] == "--each":
        parts = parts[1:]
        dn = next((c for c in ("dispatch_id", "correlation_id", "id") if c in ix), None)
        rows = [r for r in cur.execute(f"select * from {view}") if _keep(r[ix[kn]], parts)]
        if dn:
            rows.sort(key=lambda r: r[ix[dn]])
        print(f"{'dispatch':>10} {'value':>18}  counter  kernel")
        for r in rows:
            d = r[ix[dn]] if dn else "-"
            print(f"{d:>10} {float(r[ix[vn]]):18.1f}  {r[ix[cn]]}  {_short(r[ix[kn]])}")
        return
    agg = collections.defaultdict(list)
    for r in cur.execute(f"select * from {view}"):
        if _keep(r[ix[kn]], parts):
            agg[(_short(r[ix[kn]]), r[ix[cn]])].append(float(r[ix[vn]]))
    print(f"{'dispatches':>10} {'mean':>18}  counter  kernel")
    for (k, c), v in sorted(agg.items()):
        print(f"{len(v):10d} {sum(v) / len(v):18.1f}  {c}  {k}")


def _transforms(rows):
    """(first, last) row indices of each transform: from its first kernel (hash_clear_kernel,
    or fold_slab_kernel where the slab fold does the clear) to its estimate_kernel."""
    out, i = [], 0
    while i < len(rows):
        if "hash_clear_kernel" in rows[i]["name"] or "fold_slab_kernel" in rows[i]["name"]:
            j = i
            while j < len(rows) and "estimate_kernel" not in rows[j]["name"]:
                j += 1
            if j < len(rows):
                out.append((i, j))
            i = j + 1
        else:
            i += 1
    return out


def timeline(path, last):
    cur = sqlite3.connect(path).cursor()
    view = _view(cur, ("kernels",))
    ix = _columns(cur, view)
    keys = ("name", "start", "end", "grid_x", "grid_y", "workgroup_x", "lds_size", "vgpr_count")
    rows = sorted(({k: r[ix[k]] for k in keys if k in ix} for r in cur.execute(
        f"select * from {view}")), key=lambda d: d["start"])
    spans = {}
    print(f"{'precision':9s} {'span_us':>8s} {'busy_us':>8s} {'idle_us':>8s} {'kernels':>7s}")
    for i, j in _transforms(rows):
        s0, e0 = rows[i]["start"], rows[j]["end"]
        busy, cs, ce = 0, None, None
        for d in sorted(rows[i:j + 1], key=lambda d: d["start"]):
            s, e = max(d["start"], s0), min(d["end"], e0)
            if cs is None:
                cs, ce = s, e
            elif s <= ce:
                ce = max(ce, e)
            else:
                busy, cs, ce = busy + ce - cs, s, e
        busy += ce - cs
        prec = "fp32" if any("<float" in d["name"] or "unsigned int>" in d["name"]
                             for d in rows[i:j + 1]) else "fp64"
        spans.setdefault(prec, []).append((i, j))
        print(f"{prec:9s} {(e0 - s0) / 1e3:8.1f} {busy / 1e3:8.1f} {(e0 - s0 - busy) / 1e3:8.1f} "
              f"{j - i + 1:7d}")
    for prec, lst in spans.items():
        for i, j in lst[-last:]:
            s0 = rows[i]["start"]
            print(f"\n{prec}: transform starting at dispatch {i}")
            print(f"{'start_us':>8s} {'dur_us':>7s} {'grid':>12s} {'wg':>5s} {'lds_B':>6s} "
                  f"{'vgpr':>4s}  kernel")
            for d in rows[i:j + 1]:
                grid = f"{d.get('grid_x', 0)}x{d.get('grid_y', 1)}"
                print(f"{(d['start'] - s0) / 1e3:8.1f} {(d['end'] - d['start']) / 1e3:7.1f} "
                      f"{grid:>12s} {d.get('workgroup_x', 0):5d} {d.get('lds_size', 0):6d} "
                      f"{d.get('vgpr_count', 0):4d}  {_short(d['name']).replace('void ', '')}")


def _read_table(path, cols):
    """Rows of a `kernels` (cols=6) or `counters` (cols=4) text output."""
    out = []
    with open(path) as fh:
        for line in fh:
            p = line.split(None, cols - 1)
            if len(p) == cols and p[0].isdigit():
                out.append(p)
    return out


def join(kernels_txt, counter_txts):
    times = {p[5].strip(): (int(p[0]), float(p[1])) for p in _read_table(kernels_txt, 6)}
    cnt = collections.defaultdict(dict)
    for path in counter_txts:
        for p in _read_table(path, 4):
            cnt[p[3].strip()][p[2]] = float(p[1])
    print(f"{'kernel':56s} {'us':>7s} {'read MB':>8s} {'write MB':>8s} {'GB/s':>6s} "
          f"{'%peak':>5s} {'LDS conflict us/CU':>18s} {'wait %':>6s}")
    for name, (_, us) in sorted(times.items(), key=lambda kv: -kv[1][1]):
        c = cnt.get(name)
        if not c or "FETCH_SIZE" not in c:
            continue
        rd = 2 * c["FETCH_SIZE"] * 1024 / 1e6
        wr = c.get("WRITE_SIZE", 0.0) * 1024 / 1e6
        gbs = (rd + wr) / us * 1e3
        conf = c.get("SQ_LDS_BANK_CONFLICT", 0.0) / 256 / 2400.0
        wait = 100.0 * c.get("SQ_WAIT_ANY", 0.0) / max(1.0, c.get("SQ_WAVE_CYCLES", 1.0))
        short = name.replace("void ", "").replace("sfft::", "")[:56]
        print(f"{short:56s} {us:7.1f} {rd:8.1f} {wr:8.1f} {gbs:6.0f} {100 * gbs / 8000:5.1f} "
              f"{conf:18.1f} {wait:6.1f}")


def main(argv):
    if len(argv) >= 4 and argv[1] == "join":
        return join(argv[2], argv[3:])
    if len(argv) >= 3 and argv[1] == "timeline":
        return timeline(argv[2], int(argv[3]) if len(argv) > 3 else 1)
    if len(argv) < 3 or argv[1] not in ("kernels", "counters"):
        raise SystemExit(__doc__)
    (kernels if argv[1] == "kernels" else counters)(argv[2], argv[3:])


if __name__ == "__main__":
    main(sys.argv)
