"""Regenerate the golden fixtures from the compiled reference `comp`.

Runs in the development container only (it needs oracle/_ref, built by
`make -C oracle ref` from the reference sources). For every case it writes a
ParallelMETA directory (db.config, tree/id.txt, tree/order.txt,
copy_number.txt) and an OTU table with this repository's generator, runs the
reference

    comp -T table -o out.dist -D G -M 0 [-r F] [-d F]       (oracle/_ref/comp_O0)
    ref_driver --hex (_Comp_Tree::Calc_sim of every pair)  (oracle/_ref/ref_driver_O0)

and stores inputs + outputs as tests/golden/<case>.tar.gz:

    input/databases/...      the database as the reference read it
    input/table.txt          the -T table
    dist_rT.dist             comp -r T (copy-number correction on)
    dist_rF.dist             comp -r F
    sim_rT.dist              comp -r T -d F (similarities)
    dist_M1.dist             comp -M 1 (unweighted Meta-Storms)
    hex_rT.txt / hex_rF.txt  "i j <C99 hex float>" per pair, Calc_sim FP32
    hex_M1.txt               the same for Calc_sim_unweight (comp.h:378-446)
    hex_M2..M5.txt           the same for Calc_sim(a, b, mode) modes 2-5
                             (comp.h:296-330: cosine, Euclidean, JSD, Bray-Curtis)

The `files` case holds the single-file inputs of `-i` and `-l` instead of a
table: classification files in both formats (count lines and per-read
lines) with the loader's quirks, and three sample lists; it stores

    abd_rT.txt / abd_rF.txt  ref_driver --files: Load_abd(file) per file,
                             "# k mapped" then "leaf <hex>" per nonzero leaf
    l_<list>_<tag>.dist      comp -l <list> [-p samples/] ... outputs
    i_<tag>.txt              comp -i file1 file2 ... standard output

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import importlib
import io
import os
import subprocess
import sys
import tarfile
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
cms = importlib.import_module("coordinated-meta-storms_amd")
REF = os.path.join(ROOT, "oracle", "_ref")


def toy_case(d):
    """4-leaf tree ((A:0.1,B:0.2):0.3,(C:0.4,D:0.5):0.6), copy number A = 2."""
    dbdir = os.path.join(d, "databases", "toy", "tree")
    os.makedirs(dbdir)
    with open(os.path.join(d, "databases", "db.config"), "w") as f:
        f.write("G\n0\ntoy\nToy tree\nN\nY\nY\nN\n")
    with open(os.path.join(dbdir, "id.txt"), "w") as f:
        f.write("A\nB\nC\nD\n")
    with open(os.path.join(dbdir, "order.txt"), "w") as f:
        f.write("0 0.1 1 0.2 -1\n2 0.4 3 0.5 -2\n-1 0.3 -2 0.6 -1\n")
    with open(os.path.join(d, "databases", "toy", "copy_number.txt"), "w") as f:
        f.write("ID\tcp\nA\t2\nB\t1\nC\t1\nD\t1\n")
    with open(os.path.join(d, "table.txt"), "w") as f:
        f.write("SampleID\tA\tB\tC\tD\ns1\t1\t1\t0\t0\ns2\t1\t0\t1\t0\ns3\t0\t1\t1\t0\n")


def quirks_case(d):
    """200-leaf synthetic tree, 12 samples, with the loader's quirks in the
    table: an OTU_ prefix (stripped), an otu_ prefix, an unmapped column, a
    duplicated OTU column (counted twice in the total, stored once), and one
    row of the wrong width (skipped)."""
    tree = cms.synth_tree(4242, 200, 0.05, 0.99)
    cp = cms.synth_copy_numbers(4243, 200, 15)
    cms.write_db(d, "quirks", tree, cp)
    counts = cms.synth_counts(4244, 12, 200, 10, 60, 1.1, 50, 0.5)
    dense = counts.dense(200)
    names = [f"T{i}" for i in range(200)]
    names[3] = "OTU_T3"
    names[7] = "otu_T7"
    extra_names = ["unmapped_otu", "T11"]  # T11 appears twice
    extra = np.stack([dense[:, 5] + 3, dense[:, 11] + 2], axis=1)
    full = np.concatenate([dense, extra], axis=1)
    with open(os.path.join(d, "table.txt"), "w") as f:
        f.write("SampleID\t" + "\t".join(names + extra_names) + "\n")
        for i in range(full.shape[0]):
            f.write(f"Q{i}\t" + "\t".join("%g" % v for v in full[i]) + "\n")
            if i == 4:
                f.write("BROKEN\t" + "\t".join("1" for _ in range(full.shape[1] - 1)) + "\n")


def config0_case(d):
    """Workload 0: 5,000-leaf Yule tree, 64 samples (seeds 1000/2000/3000)."""
    tree, cp, counts, _ = cms.workload(0)
    cms.write_db(d, "cfg0", tree, cp)
    cms.write_table(os.path.join(d, "table.txt"), counts, tree.leaf_n)


def core_case(d):
    """The reference's only real tree, databases/oral_core/tree/core.tre
    (1,045 leaves; copied to tests/golden/core.tre), turned into a database
    by this repository's Newick builder; synthetic copy numbers and 40
    samples over its leaf ids (seeds 4342/4343)."""
    nw = cms.Newick(os.path.join(HERE, "core.tre"))
    names = nw.names
    cp = cms.synth_copy_numbers(4342, len(names), 15)
    cp_path = os.path.join(d, "cp_in.txt")
    with open(cp_path, "w") as f:
        f.write("ID\tcopy_number\n" + "".join(f"{n}\t{int(c)}\n" for n, c in zip(names, cp)))
    nw.write_db(d, "oral_core", cp_path)
    os.remove(cp_path)
    counts = cms.synth_counts(4343, 40, len(names), 20, 300, 1.1, 50, 0.5)
    dense = counts.dense(len(names))
    with open(os.path.join(d, "table.txt"), "w") as f:
        f.write("SampleID\t" + "\t".join(names) + "\n")
        for i in range(dense.shape[0]):
            f.write(f"C{i}\t" + "\t".join("%g" % v for v in dense[i]) + "\n")


CASES = {"toy": toy_case, "quirks": quirks_case, "config0": config0_case, "core": core_case}

# -i / -l inputs: (list file, -p prefix or None, outputs as (tag, extra args))
FILE_LISTS = {
    "ids": (None, [("M0", ["-M", "0"]), ("rF", ["-M", "0", "-r", "F"]), ("M1", ["-M", "1"]),
                   ("M4", ["-M", "4"])]),
    "paths": (None, [("sim", ["-M", "0", "-d", "F"])]),
    "prefix": ("samples/", [("M0", ["-M", "0"]), ("M5", ["-M", "5"])]),
}
# -i pairs: (tag, file1, file2, extra args)
FILE_PAIRS = [
    ("M0", "samples/S0/classification.txt", "samples/S1/classification.txt", ["-M", "0"]),
    ("sim_rF", "samples/S2/classification.txt", "samples/R0/classification.txt", ["-M", "0", "-d", "F", "-r", "F"]),
    ("M1", "samples/S3/classification.txt", "samples/R1/classification.txt", ["-M", "1"]),
    ("M4", "samples/S4/classification.txt", "samples/S5/classification.txt", ["-M", "4"]),
    ("M5", "samples/R2/classification.txt", "samples/S6/classification.txt", ["-M", "5"]),
]


def files_case(d):
    """600-leaf synthetic tree; 9 count-format files S0..S8 and 3 per-read
    files R0..R2 under samples/<id>/classification.txt. The count files carry
    the loader's quirks: OTUs outside the tree (they count in the total), an
    id given twice (the last count wins), zero counts, an id without a count,
    counts "12.7" and "3e2" (read as 12 and 3), an unstripped OTU_ prefix,
    leading blanks, CRLF line ends, and more than 193 distinct ids (the hash
    map grows while loading, which changes the order of the FP32 total).
    N0 has a negative count (loader check only)."""
    rng = np.random.default_rng(4444)
    tree = cms.synth_tree(4442, 600, 0.05, 0.99)
    cp = cms.synth_copy_numbers(4443, 600, 15)
    cms.write_db(d, "files", tree, cp)
    paths = []

    def put(name, text):
        p = os.path.join(d, "samples", name)
        os.makedirs(p, exist_ok=True)
        with open(os.path.join(p, "classification.txt"), "w", newline="") as f:
            f.write(text)
        paths.append(f"samples/{name}/classification.txt")

    for k in range(9):
        n_ids = [40, 150, 260, 420, 90, 330, 500, 210, 60][k]
        leaves = rng.choice(600, size=min(n_ids, 560), replace=False)
        lines = ["#Database_OTU\tCount\tAbundance\tTaxonomy"]
        for l in leaves:
            c = int(rng.negative_binomial(0.5, 0.5 / 50.5)) + 1
            lines.append(f"T{l}\t{c}\t0.0\tk__Bacteria; otu_T{l}")
        for j in range(int(rng.integers(0, 80))):
            lines.append(f"U{k}_{j}\t{int(rng.integers(1, 40))}\t0.0\tUnclassified")
        if k % 3 == 1:
            lines.append(f"T{leaves[0]}\t{int(rng.integers(1, 99))}\t0\tdup")      # last count wins
            lines.append(f"T{leaves[1]}\t0\t0\tzero")
            lines.append(f"T{leaves[2]}")                                            # no count: 0
            lines.append(f"OTU_T{leaves[3]}\t7\t0\tprefixed")                        # not stripped
        if k % 3 == 2:
            lines.append(f"  T{leaves[4]}\t12.7\t0\tblank-led")
            lines.append(f"T{leaves[5]}\t3e2\t0\texp")
        eol = "\r\n" if k == 5 else "\n"
        put(f"S{k}", eol.join(lines) + eol)
    for k in range(3):
        lines = ["#Seq_ID\tDatabase_OTU\tSimilarity"]
        leaves = rng.choice(600, size=[30, 120, 240][k], replace=False)
        for r in range([400, 900, 1500][k]):
            l = leaves[min(int(rng.zipf(1.3)) - 1, len(leaves) - 1)]
            lines.append(f"read{r}\tT{l}\t0.97")
        put(f"R{k}", "\n".join(lines) + "\n")
    put("N0", "#Database_OTU\tCount\nT1\t5\nT2\t-3\nT3\t9\n")
    with open(os.path.join(d, "files_all.txt"), "w") as f:
        f.write("\n".join(paths) + "\n")
    with open(os.path.join(d, "list_ids.txt"), "w") as f:
        for k in (0, 1, 2, 3, 4, 5, 6, 7, 8):
            f.write(f"id{k}\tsamples/S{k}/classification.txt\n")
        for k in range(3):
            f.write(f"rd{k} samples/R{k}/classification.txt extra\n")
    with open(os.path.join(d, "list_paths.txt"), "w") as f:
        for k in (3, 1, 4, 1, 5):
            f.write(f"samples/S{k}/classification.txt\n")
        f.write("named samples/R0/classification.txt\n")
        f.write("./samples/S8/classification.txt")                                  # no final newline
    with open(os.path.join(d, "list_prefix.txt"), "w") as f:
        for k in (2, 7, 0):
            f.write(f"S{k}/classification.txt\n")
        f.write("\n")                                                               # empty line: skipped
        f.write("R2/classification.txt\n")


def run_ref(d, args, out=None):
    env = dict(os.environ, ParallelMETA=d)
    r = subprocess.run(args, cwd=d, env=env, capture_output=True, text=True, check=True)
    return r.stdout


def make_files():
    with tempfile.TemporaryDirectory() as tmp:
        d = os.path.join(tmp, "input")
        os.makedirs(d)
        files_case(d)
        comp = os.path.join(REF, "comp_O0")
        drv = os.path.join(REF, "ref_driver_O0")
        outs = {}
        for tag, extra in [("rT", []), ("rF", ["-r", "F"])]:
            outs[f"abd_{tag}.txt"] = run_ref(d, [drv, "-D", "G", "--files", "files_all.txt"] + extra).encode()
        for name, (prefix, runs) in FILE_LISTS.items():
            for tag, extra in runs:
                o = os.path.join(tmp, f"l_{name}_{tag}.dist")
                args = [comp, "-l", f"list_{name}.txt", "-o", o, "-D", "G"] + extra
                if prefix:
                    args += ["-p", prefix]
                run_ref(d, args)
                outs[f"l_{name}_{tag}.dist"] = open(o, "rb").read()
        for tag, f1, f2, extra in FILE_PAIRS:
            outs[f"i_{tag}.txt"] = run_ref(d, [comp, "-i", f1, f2, "-D", "G"] + extra).encode()
        write_case("files", d, outs)


def write_case(case, d, outs):
    path = os.path.join(HERE, case + ".tar.gz")
    with tarfile.open(path, "w:gz") as tf:
        tf.add(d, arcname="input")
        for name, data in outs.items():
            ti = tarfile.TarInfo(name)
            ti.size = len(data)
            tf.addfile(ti, io.BytesIO(data))
    print(f"{path}: {os.path.getsize(path)} bytes")


def make(case):
    if case == "files":
        make_files()
        return
    with tempfile.TemporaryDirectory() as tmp:
        d = os.path.join(tmp, "input")
        os.makedirs(d)
        CASES[case](d)
        comp = os.path.join(REF, "comp_O0")
        drv = os.path.join(REF, "ref_driver_O0")
        outs = {}
        for tag, extra in [("dist_rT", []), ("dist_rF", ["-r", "F"]), ("sim_rT", ["-d", "F"]),
                           ("dist_M1", ["-M", "1"])]:
            o = os.path.join(tmp, tag + ".dist")
            args = [comp, "-T", "table.txt", "-o", o, "-D", "G"] + (extra if "-M" in extra else ["-M", "0"] + extra)
            run_ref(d, args)
            outs[tag + ".dist"] = open(o, "rb").read()
        outs["hex_rT.txt"] = run_ref(d, [drv, "-D", "G", "-T", "table.txt", "--hex"]).encode()
        outs["hex_rF.txt"] = run_ref(d, [drv, "-D", "G", "-T", "table.txt", "-r", "F", "--hex"]).encode()
        outs["hex_M1.txt"] = run_ref(d, [drv, "-D", "G", "-T", "table.txt", "-M", "1", "--hex"]).encode()
        for m in (2, 3, 4, 5):  # cosine, Euclidean, JSD, Bray-Curtis (dist.h)
            outs[f"hex_M{m}.txt"] = run_ref(d, [drv, "-D", "G", "-T", "table.txt", "-M", str(m), "--hex"]).encode()
        write_case(case, d, outs)


if __name__ == "__main__":
    for c in (sys.argv[1:] or list(CASES) + ["files"]):
        make(c)
