"""Summaries of rocprofv3 output.

Kernel table (calls, total / average / min / max duration, share) of a
kernel trace, read from its SQLite results database (`rocprofv3
--kernel-trace -d OUT -o NAME`, file OUT/NAME_results.db) or from a
`--output-format csv` NAME_kernel_trace.csv:

    python3 lightmetrica-v3_amd/lm_amd/prof_summary.py OUT/NAME_kernel_trace.csv [--match SUBSTR] [--top N]

Counter summary of one kernel (JSON) from the --pmc passes of
lm_amd/profile_trace.py (FETCH_SIZE, WRITE_SIZE, TCC_HIT/MISS, SQ_*):

    python3 lightmetrica-v3_amd/lm_amd/prof_summary.py --pmc 'k_trace<4' --fetch F.csv --write W.csv \
        [--tcc T.csv] [--sq S.csv] [--trace KT.csv]
"""
import argparse
import csv
import re
import sqlite3
import sys
from collections import defaultdict


def load(path):
    rows = []
    if path.endswith('.db'):
        db = sqlite3.connect(path)
        for name, dur in db.execute('select name, duration from kernels'):
            rows.append((name, float(dur)))
    else:
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((r['Kernel_Name'], float(r['End_Timestamp']) - float(r['Start_Timestamp'])))
    return rows


def short(name: str) -> str:
    name = name.replace('(anonymous namespace)::', '')
    name = re.sub(r'\(.*$', '', name)
    name = re.sub(r'^void ', '', name)
    return name[:70]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('path', nargs='?')
    ap.add_argument('--match', default='')
    ap.add_argument('--top', type=int, default=30)
    ap.add_argument('--pmc', default='', help='kernel name substring: print its counter summary (JSON)')
    ap.add_argument('--fetch')
    ap.add_argument('--write')
    ap.add_argument('--tcc')
    ap.add_argument('--sq')
    ap.add_argument('--trace')
    ap.add_argument('--units', type=int, default=0)
    a = ap.parse_args(argv)
    if a.pmc:
        import json
        print(json.dumps(pmc_summary(a.pmc, a.fetch, a.write, a.tcc, a.sq, a.trace or a.path,
                                     units_per_launch=a.units or None), indent=1))
        return 0
    stats = defaultdict(list)
    for name, dur in load(a.path):
        if a.match in name:
            stats[short(name)].append(dur)
    total = sum(sum(v) for v in stats.values()) or 1.0
    print('| kernel | calls | total us | avg us | min us | max us | % |')
    print('|---|---|---|---|---|---|---|')
    for k, v in sorted(stats.items(), key=lambda kv: -sum(kv[1]))[:a.top]:
        s = sum(v)
        print(f'| {k} | {len(v)} | {s / 1e3:.1f} | {s / len(v) / 1e3:.1f} | {min(v) / 1e3:.1f} | '
              f'{max(v) / 1e3:.1f} | {100 * s / total:.2f} |')



def pmc_summary(kernel_substr, fetch_csv, write_csv, tcc_csv=None, sq_csv=None, trace_csv=None, copy_bytes=512 << 20,
                units_per_launch=None):
    """HBM traffic and efficiency counters of one kernel from rocprofv3
    --pmc runs of lm_amd/profile_trace.py. FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch) are scaled by what they read for the profile driver's device
    copy of `copy_bytes` bytes (reads and writes copy_bytes)."""
    def per_dispatch(path, match):
        out = defaultdict(dict)
        with open(path) as f:
            for r in csv.DictReader(f):
                if match(r['Kernel_Name']):
                    out[int(r['Dispatch_Id'])][r['Counter_Name']] = float(r['Counter_Value'])
        return out

    def mean(d, name):
        v = [x[name] for x in d.values() if name in x]
        return sum(v) / len(v) if v else None

    kern = lambda n: kernel_substr in n
    copy = lambda n: 'copyBuffer' in n
    fk = mean(per_dispatch(fetch_csv, kern), 'FETCH_SIZE')
    wk = mean(per_dispatch(write_csv, kern), 'WRITE_SIZE')
    fc = mean(per_dispatch(fetch_csv, copy), 'FETCH_SIZE')
    wc = mean(per_dispatch(write_csv, copy), 'WRITE_SIZE')
    # the largest copy is the calibration copy
    fcs = [x['FETCH_SIZE'] for x in per_dispatch(fetch_csv, copy).values()]
    wcs = [x['WRITE_SIZE'] for x in per_dispatch(write_csv, copy).values()]
    fc, wc = max(fcs), max(wcs)
    fscale = copy_bytes / (fc * 1024.0)
    wscale = copy_bytes / (wc * 1024.0)
    out = {'kernel': kernel_substr, 'FETCH_SIZE_KiB': fk, 'WRITE_SIZE_KiB': wk,
           'copy_calibration': {'bytes': copy_bytes, 'FETCH_SIZE_KiB': fc, 'WRITE_SIZE_KiB': wc,
                                'fetch_scale': fscale, 'write_scale': wscale},
           'hbm_bytes_per_launch': fk * 1024.0 * fscale + wk * 1024.0 * wscale}
    if trace_csv:
        durs = [d for n, d in load(trace_csv) if kernel_substr in n]
        out['avg_kernel_ns'] = sum(durs) / len(durs)
        out['hbm_GBs'] = out['hbm_bytes_per_launch'] / out['avg_kernel_ns']
    if units_per_launch:
        out['units_per_launch'] = units_per_launch
    if tcc_csv:
        t = per_dispatch(tcc_csv, kern)
        hit, miss = mean(t, 'TCC_HIT_sum'), mean(t, 'TCC_MISS_sum')
        out['TCC_HIT_sum'], out['TCC_MISS_sum'] = hit, miss
        out['l2_hit_rate'] = hit / (hit + miss) if hit is not None and miss is not None else None
    if sq_csv:
        sq = per_dispatch(sq_csv, kern)
        names = sorted({k for d in sq.values() for k in d})
        out['SQ'] = {n: mean(sq, n) for n in names}
        if 'SQ_THREAD_CYCLES_VALU' in out['SQ'] and 'SQ_ACTIVE_INST_VALU' in out['SQ']:
            # average active lanes per VALU instruction (of 64)
            out['valu_active_lanes'] = out['SQ']['SQ_THREAD_CYCLES_VALU'] / out['SQ']['SQ_ACTIVE_INST_VALU']
    return out


if __name__ == '__main__':
    sys.exit(main())
