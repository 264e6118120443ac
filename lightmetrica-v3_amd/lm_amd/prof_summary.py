"""Per-kernel summary (calls, total / average / min / max duration, share)
of a rocprofv3 kernel trace, read from its SQLite results database
(`rocprofv3 --kernel-trace -d OUT -o NAME`, file OUT/NAME_results.db) or
from a `--output-format csv` kernel_trace.csv. Prints a markdown table.

    python3 lightmetrica-v3_amd/lm_amd/prof_summary.py OUT/NAME_results.db [--match SUBSTR] [--top N]
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
    ap.add_argument('path')
    ap.add_argument('--match', default='')
    ap.add_argument('--top', type=int, default=30)
    a = ap.parse_args(argv)
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


if __name__ == '__main__':
    sys.exit(main())
