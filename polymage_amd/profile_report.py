"""Summaries of rocprofv3 output for the repository's ``profiles/`` folder.

    python -m polymage_amd.profile_report --kernel pmx_unsharp_g0 \
        --trace DIR --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR \
        --sq DIR --copy-mib 512 --workload unsharp --out profiles/NAME.md \
        --traffic-json profiles/traffic.json --code-key KEY --ideal BYTES

Reads the CSV files rocprofv3 writes with ``--output-format csv``
(``*_kernel_stats.csv`` of a ``--kernel-trace --stats`` run and
``*_counter_collection.csv`` of ``--pmc`` runs) and writes a short markdown
summary: kernel statistics, FETCH_SIZE / WRITE_SIZE per dispatch scaled by
what the same counters read for a plain device copy of known size, the HBM
traffic against the ideal (input + output) bytes, and SQ instruction
counters.  Optionally records the measured traffic for ``bench.py``.
"""

import argparse
import csv
import glob
import json
import os
import statistics


def _csv(d, suffix):
    files = sorted(glob.glob(os.path.join(d, '*' + suffix)))
    if not files:
        raise FileNotFoundError('%s/*%s' % (d, suffix))
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def counter(d, kernel, name):
    """Per-dispatch values of counter ``name`` for kernels whose name starts
    with ``kernel`` (None: the largest-valued kernel of the run)."""
    rows = [r for r in _csv(d, '_counter_collection.csv') if r['Counter_Name'] == name]
    if kernel:
        rows = [r for r in rows if r['Kernel_Name'].startswith(kernel)]
    else:
        best = max(rows, key=lambda r: float(r['Counter_Value']))['Kernel_Name']
        rows = [r for r in rows if r['Kernel_Name'] == best]
    return [float(r['Counter_Value']) for r in rows]


def kernel_means(d, prefix, name):
    """{kernel: mean per dispatch of counter ``name``} for the kernels whose
    name starts with ``prefix``."""
    vals = {}
    for r in _csv(d, '_counter_collection.csv'):
        k = r['Kernel_Name'].split('(')[0]
        if r['Counter_Name'] == name and k.startswith(prefix):
            vals.setdefault(k, []).append(float(r['Counter_Value']))
    return dict((k, statistics.mean(v)) for k, v in vals.items())


SQ_NAMES = ['SQ_WAVES', 'SQ_INSTS_VALU', 'SQ_ACTIVE_INST_VALU', 'SQ_INSTS_SALU', 'SQ_INSTS_SMEM',
            'SQ_INSTS_VMEM_RD', 'SQ_INSTS_VMEM_WR', 'SQ_INSTS_LDS', 'SQ_LDS_BANK_CONFLICT',
            'SQ_BUSY_CYCLES', 'SQ_WAVE_CYCLES', 'SQ_WAIT_INST_ANY', 'GRBM_GUI_ACTIVE']
# CUs of one MI355X and its peak engine clock: a CU issues at most one
# wave64 VALU instruction per cycle (four 16-lane SIMDs, four cycles each)
N_CU = 256
CLOCK_MHZ = 2400.0


def sq_table(dirs, prefix, kernel_us):
    """Per-kernel means per dispatch of the SQ/GRBM counters in the --pmc
    runs ``dirs``, for kernels whose name starts with ``prefix``, with two
    derived columns: VALU busy (SQ_INSTS_VALU / (CUs x cycles of the traced
    average duration at the peak clock)) and LDS bank-conflict cycles per
    LDS instruction."""
    vals = {}
    for d in dirs:
        for r in _csv(d, '_counter_collection.csv'):
            k = r['Kernel_Name'].split('(')[0]
            if not k.startswith(prefix) or r['Counter_Name'] not in SQ_NAMES:
                continue
            vals.setdefault(k, {}).setdefault(r['Counter_Name'], []).append(float(r['Counter_Value']))
    if not vals:
        return []
    have = [n for n in SQ_NAMES if any(n in v for v in vals.values())]
    L = ['SQ counters, mean per dispatch (`--pmc`, counters summed over the GPU):', '',
         '| kernel | us | ' + ' | '.join(n.replace('SQ_', '') for n in have) +
         ' | VALU busy | conflict cycles / LDS instr |',
         '|---|---|' + '---|' * len(have) + '---|---|']
    for k in sorted(vals):
        v = dict((n, statistics.mean(x)) for n, x in vals[k].items())
        us = kernel_us.get(k)
        busy = (v['SQ_INSTS_VALU'] / (N_CU * us * CLOCK_MHZ)) if (us and 'SQ_INSTS_VALU' in v) else None
        conf = (v['SQ_LDS_BANK_CONFLICT'] / v['SQ_INSTS_LDS']) if v.get('SQ_INSTS_LDS') else None
        L.append('| %s | %s | %s | %s | %s |' % (
            k, '%.2f' % us if us else '-', ' | '.join('%.4g' % v[n] if n in v else '-' for n in have),
            '%.2f' % busy if busy is not None else '-', '%.3f' % conf if conf is not None else '-'))
    L += ['', 'VALU busy = SQ_INSTS_VALU / (%d CUs x traced average duration x %.0f MHz): a CU '
              'issues at most one wave64 VALU instruction per cycle.' % (N_CU, CLOCK_MHZ), '']
    return L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--kernel', required=True)
    ap.add_argument('--workload', required=True)
    ap.add_argument('--trace')
    ap.add_argument('--fetch')
    ap.add_argument('--write')
    ap.add_argument('--copy-fetch')
    ap.add_argument('--copy-write')
    ap.add_argument('--copy-mib', type=float, default=512)
    ap.add_argument('--sq', action='append', default=[],
                    help='a --pmc run of SQ/GRBM counters (may be repeated)')
    ap.add_argument('--ideal', type=float, default=None, help='ideal bytes per dispatch')
    ap.add_argument('--ideal-in', type=float, default=None, help='input bytes per dispatch')
    ap.add_argument('--commands', default='', help='text: the commands that made the data')
    ap.add_argument('--calls', type=int, default=0,
                    help='pipelines of several kernels: traffic per pipeline call; N > 0 sums '
                         'the matching dispatches over N calls, -1 sums each matching kernel\'s '
                         'mean per dispatch (every kernel runs once per call); 0 = average per dispatch')
    ap.add_argument('--out', required=True)
    ap.add_argument('--traffic-json')
    ap.add_argument('--code-key')
    ap.add_argument('--traffic-key', default=None,
                    help='record name in --traffic-json (default: --workload)')
    args = ap.parse_args()

    L = ['# %s — %s' % (args.kernel, args.workload), '']
    if args.commands:
        L += ['Commands (repository root, one MI355X):', '', '```', args.commands.strip(), '```', '']
    avg_us = None
    kernel_us = {}
    if args.trace:
        rows = _csv(args.trace, '_kernel_stats.csv')
        for r in rows:
            kernel_us[r['Name']] = float(r['AverageNs']) / 1e3
        L += ['Kernel statistics (`--kernel-trace --stats`):', '',
              '| kernel | calls | average (us) | min (us) | max (us) | % |', '|---|---|---|---|---|---|']
        per_call = 0.0
        for r in rows:
            L.append('| %s | %s | %.2f | %.2f | %.2f | %.2f |' % (
                r['Name'][:60], r['Calls'], float(r['AverageNs']) / 1e3, float(r['MinNs']) / 1e3,
                float(r['MaxNs']) / 1e3, float(r['Percentage'])))
            if r['Name'] == args.kernel or (not args.calls and r['Name'].startswith(args.kernel)):
                avg_us = float(r['AverageNs']) / 1e3
            elif args.calls and r['Name'].startswith(args.kernel):
                # every kernel of the pipeline runs once per call
                per_call += float(r['AverageNs']) / 1e3
        if avg_us is None and per_call:
            avg_us = per_call
            L += ['', 'Sum of the average durations of the `%s*` kernels (one pipeline call): '
                      '%.2f us.' % (args.kernel, per_call)]
        L.append('')
    traffic = None
    if args.fetch and args.write and args.copy_fetch and args.copy_write:
        if args.calls < 0:
            # multi-kernel pipeline: each kernel's mean per dispatch, summed
            f = sum(kernel_means(args.fetch, args.kernel, 'FETCH_SIZE').values())
            wr = sum(kernel_means(args.write, args.kernel, 'WRITE_SIZE').values())
        elif args.calls:
            # multi-kernel pipeline: all matching dispatches per pipeline call
            f = sum(counter(args.fetch, args.kernel, 'FETCH_SIZE')) / args.calls
            wr = sum(counter(args.write, args.kernel, 'WRITE_SIZE')) / args.calls
        else:
            f = statistics.mean(counter(args.fetch, args.kernel, 'FETCH_SIZE'))
            wr = statistics.mean(counter(args.write, args.kernel, 'WRITE_SIZE'))
        cf = statistics.median(counter(args.copy_fetch, '__amd_rocclr_copyBuffer', 'FETCH_SIZE'))
        cw = statistics.median(counter(args.copy_write, '__amd_rocclr_copyBuffer', 'WRITE_SIZE'))
        kib = args.copy_mib * 1024
        ff, fw = kib / cf, kib / cw
        fetch_b = f * ff * 1024
        write_b = wr * fw * 1024
        traffic = fetch_b + write_b
        L += ['HBM traffic per dispatch (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`, separate runs):', '',
              '| | counter (KiB) | copy-probe factor | bytes |', '|---|---|---|---|',
              '| FETCH_SIZE | %.1f | %.3f | %.0f |' % (f, ff, fetch_b),
              '| WRITE_SIZE | %.1f | %.3f | %.0f |' % (wr, fw, write_b),
              '| total | | | %.0f |' % traffic, '',
              'The copy probe is a device-to-device copy of %g MiB; the counters read %.1f KiB '
              '(FETCH) and %.1f KiB (WRITE) for it, hence the factors.' % (args.copy_mib, cf, cw)]
        if args.ideal:
            L.append('Ideal traffic (input + output) %.0f bytes: measured / ideal = %.3f.' % (
                args.ideal, traffic / args.ideal))
        if args.ideal_in:
            L.append('Reads / input bytes = %.3f; writes / output bytes = %.3f.' % (
                fetch_b / args.ideal_in, write_b / max(1.0, args.ideal - args.ideal_in)))
        if avg_us:
            L.append('At the traced average of %.2f us: ideal %.1f GB/s, measured traffic %.1f GB/s.' % (
                avg_us, (args.ideal or 0) / avg_us / 1e3, traffic / avg_us / 1e3))
        L.append('')
    if args.sq:
        L += sq_table(args.sq, args.kernel, kernel_us)
    with open(args.out, 'w') as fo:
        fo.write('\n'.join(L) + '\n')
    if args.traffic_json and traffic is not None:
        rec = {}
        if os.path.exists(args.traffic_json):
            with open(args.traffic_json) as fi:
                rec = json.load(fi)
        rec[args.traffic_key or args.workload] = {'bytes_per_dispatch': round(traffic), 'code_key': args.code_key,
                              'per': 'pipeline call' if args.calls else 'dispatch',
                              'kernel': args.kernel, 'summary': os.path.basename(args.out)}
        with open(args.traffic_json, 'w') as fo:
            json.dump(rec, fo, indent=1)
    print('\n'.join(L))


if __name__ == '__main__':
    main()
