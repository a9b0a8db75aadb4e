"""Reader for the reference's .ref sample files (file_input.cpp:123-284 layout), used by the
tests to pin the oracle against Whizard's own momenta."""
from __future__ import annotations

import numpy as np

from gpu_phs_amd.process import Process


def read_ref(path):
    with open(path) as f:
        lines = f.read().splitlines()
    hdr = [int(l.split()[1]) for l in lines[:6]]
    n_ch, n_in, n_out, n_x = hdr[0], hdr[1], hdr[2], hdr[5]
    n_prt, n_prt_out = 2 ** (n_in + n_out) - 1, 2 ** n_out - 1
    pos = 6
    d1, d2, hc, fr, mid, mm, mw = [], [], [], [], [], [], []
    for _ in range(n_ch):
        blk = [l.split()[1:] for l in lines[pos:pos + 7]]
        pos += 7
        d1.append([int(v) for v in blk[0][:n_prt]])
        d2.append([int(v) for v in blk[1][:n_prt]])
        hc.append([int(v) for v in blk[2][:n_prt]])
        fr.append(int(blk[3][0]))
        mid.append([int(v) for v in blk[4][:n_prt_out]])
        mm.append([float(v) for v in blk[5][:n_prt_out]])
        mw.append([float(v) for v in blk[6][:n_prt_out]])
    flv = np.array([float(v) for v in lines[pos].split()[1:]])
    flw = np.array([float(v) for v in lines[pos + 1].split()[1:]])
    pos += 2
    per = 5 + n_in + n_out
    ch, xs, moms, vols, oks, facs = [], [], [], [], [], []
    while pos + per <= len(lines):
        blk = lines[pos:pos + per]
        pos += per
        ch.append(int(blk[0].split()[1]) - 1)
        xs.append([float(v) for v in blk[1].split()[1:]])
        moms.append([[float(v.replace("D", "E")) for v in b.split()] for b in blk[2:2 + n_in + n_out]])
        vols.append(float(blk[2 + n_in + n_out].split()[1]))
        oks.append(int(blk[3 + n_in + n_out].split()[1]))
        facs.append([float(v) for v in blk[4 + n_in + n_out].split()[1:]])
    sqrts = float(moms[0][0][0] + moms[0][1][0]) if moms else 0.0
    proc = Process(n_in, n_out, np.array(d1, np.int32), np.array(d2, np.int32), np.array(hc, np.int32),
                   np.array(mid, np.int32), np.array(mm), np.array(mw), flv, flw, sqrts, name=path)
    ev = dict(channel=np.array(ch, np.int32), x=np.array(xs), p=np.array(moms), volume=np.array(vols),
              ok=np.array(oks, bool), factors=facs, hasfriends=np.array(fr, np.int32))
    return proc, ev
