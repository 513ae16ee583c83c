"""Text summaries of rocprofv3 result databases (the SQLite files `rocprofv3 -d DIR -o run`
writes), used for the kernel timings and hardware counters kept under profiles/.

    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py kernels RUN_results.db [NAME_PART ...]
    python3 gpu-sfft_amd/sfft_amd/rocpd_summary.py counters RUN_results.db [NAME_PART ...]

`kernels`: per kernel name, the number of dispatches, mean / min / max duration (us) and the
share of the summed kernel time (from the `kernels` view: end - start in ns).
`counters`: per kernel name and counter, the number of dispatches and the mean value per
dispatch (from the `counters_collection` view of a `--pmc` run).
NAME_PART filters the kernel names (substring match, any of them).

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
    agg = collections.defaultdict(list)
    for r in cur.execute(f"select * from {view}"):
        if _keep(r[ix[kn]], parts):
            agg[(_short(r[ix[kn]]), r[ix[cn]])].append(float(r[ix[vn]]))
    print(f"{'dispatches':>10} {'mean':>18}  counter  kernel")
    for (k, c), v in sorted(agg.items()):
        print(f"{len(v):10d} {sum(v) / len(v):18.1f}  {c}  {k}")


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
    if len(argv) < 3 or argv[1] not in ("kernels", "counters"):
        raise SystemExit(__doc__)
    (kernels if argv[1] == "kernels" else counters)(argv[2], argv[3:])


if __name__ == "__main__":
    main(sys.argv)
