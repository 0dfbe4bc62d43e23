"""Write the golden TNF fixtures of the counting path.

Runs in the development container only (it needs /root/reference):
  1. builds the reference's own CPU counter, openmp/TNF.cpp, from its sources
     in place (oracle/Makefile target `ref` -> oracle/_ref/tnf_ref);
  2. writes a FASTA fixture that exercises what the reference's reader and
     counter must handle: multi-line records of several widths, lower case,
     N runs and IUPAC codes, CRLF line ends, blank lines, header comments,
     lengths on both sides of the 1000 / 2500 size classes and of 4-base
     boundaries;
  3. runs tnf_ref on it and keeps its TNF.bin (FP64, one row of 136 per
     contig >= 1000 bases, zero rows for [1000, 2500)) as the expected output.

Outputs (committed): tests/golden/tnf_fixture.fa, tests/golden/tnf_fixture.f64,
tests/golden/tnf_fixture.json (row -> record index map and sizes).
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import mdet_amd as M  # noqa: E402


def make_fasta(seed: int = 11) -> bytes:
    rnd = random.Random(seed)
    cfg = M.gen_config(seed=seed, n_genomes=6, len_median=3500, len_sigma=0.8, len_min=800, len_max=16000,
                       n_frac=0.01, n_run_max=30)
    lens, _ = M.gen_lengths(cfg, 96)
    seqs = [bytearray(M.gen_ascii(cfg, i, int(lens[i]))) for i in range(96)]
    # extra records pinned to the size-class and 4-base edges
    extra_lens = [999, 1000, 1001, 1003, 2499, 2500, 2501, 2502, 2503, 2504, 3]
    for k, L in enumerate(extra_lens):
        c = M.gen_config(seed=seed + 100 + k, n_genomes=1, len_median=L, len_sigma=0.0, len_min=L, len_max=L,
                         n_frac=0.0)
        seqs.append(bytearray(M.gen_ascii(c, 0, L)))
    iupac = b"RYKMSWBDHVN"
    out = bytearray()
    for r, s in enumerate(seqs):
        # lower-case stretches
        for _ in range(rnd.randint(0, 3)):
            a = rnd.randrange(len(s))
            b = min(len(s), a + rnd.randint(1, 400))
            s[a:b] = s[a:b].lower()
        # IUPAC codes and stray bytes
        for _ in range(rnd.randint(0, 4)):
            s[rnd.randrange(len(s))] = iupac[rnd.randrange(len(iupac))]
        if r % 17 == 5:
            s[rnd.randrange(len(s))] = ord("n")
        name = f"ctg{r:03d}"
        hdr = f">{name}" + (f" len={len(s)} cov=12.5" if r % 3 == 0 else "")
        eol = b"\r\n" if r % 7 == 3 else b"\n"
        width = [60, 80, 70, 1, 7, 1000000][r % 6]
        if width == 1 and len(s) > 3000:
            width = 61
        out += hdr.encode() + eol
        for p in range(0, len(s), width):
            out += bytes(s[p:p + width]) + eol
            if r % 11 == 4 and p == 0:
                out += eol  # blank line inside the record
    return bytes(out)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    ref = os.path.join(ROOT, "oracle", "_ref", "tnf_ref")
    fa = make_fasta()
    fa_path = os.path.join(HERE, "tnf_fixture.fa")
    with open(fa_path, "wb") as f:
        f.write(fa)
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([ref, "1", fa_path], cwd=td, check=True, stdout=subprocess.DEVNULL)
        tnf = np.fromfile(os.path.join(td, "TNF.bin"), np.float64)
    recs = M.fasta_read(fa)
    rows = [k for k, (name, seq) in enumerate(recs) if len(name) > 0 and len(seq) >= 1000]
    assert tnf.size == len(rows) * 136, (tnf.size, len(rows))
    tnf.tofile(os.path.join(HERE, "tnf_fixture.f64"))
    meta = {"rows": rows, "lengths": [len(recs[k][1]) for k in rows], "n_records": len(recs),
            "source": "openmp/TNF.cpp built by oracle/Makefile, run as `tnf_ref 1 tnf_fixture.fa`"}
    with open(os.path.join(HERE, "tnf_fixture.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"{len(recs)} records, {len(rows)} TNF rows, {len(fa)} bytes of FASTA")


if __name__ == "__main__":
    main()
