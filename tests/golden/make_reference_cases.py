"""Extracts the active rows of the reference's CSV case tables into fixtures.

The reference drives its system tests from one CSV per operator (reference
test/frame/csv_loader.h:68-113); those tables are the status-code and tolerance
contract of each entry point.  This script copies their data rows (comment rows
and disabled cases dropped) into tests/golden/cases/, so the tests can run them
on machines without the reference.  Run it where the reference is mounted:

    python tests/golden/make_reference_cases.py /root/reference
"""
import csv
import os
import sys

TABLES = {
    "sgetrf_batched.csv": "test/getrf_batched/sgetrf_batched/arch35/sgetrf_batched_test.csv",
    "getrs_batched.csv": "test/getrs_batched/arch35/getrs_batched_test.csv",
    "sgetri_batched.csv": "test/getri_batched/sgetri_batched/arch35/sgetri_batched_test.csv",
    "smatinv_batched.csv": "test/matinv_batched/arch35/matinv_batched_test.csv",
    "sgemm_ex.csv": "test/sgemm_ex/arch35/sgemm_ex_test.csv",
    "gemm_ex.csv": "test/gemm/gemm_ex/arch35/gemm_ex_test.csv",
    "gemm_batched_ex.csv": "test/gemm_batched_ex/arch35/gemm_batched_ex_test.csv",
    "gemm_grouped_batched.csv": "test/gemm_grouped_batched/arch35/gemm_grouped_batched_test.csv",
    "gemm_grouped_batched_ex.csv": "test/gemm_grouped_batched_ex/arch35/gemm_grouped_batched_ex_test.csv",
    "sgemm_batched.csv": "test/gemm_batched/sgemm_batched/arch35/sgemm_batched_test.csv",
    "strsm.csv": "test/trsm/strsm/arch35/strsm_test.csv",
    "blaslt_matmul.csv": "test/blasLtMatmul/arch35/blasLtMatmul_test.csv",
}


def main(ref_root):
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cases")
    os.makedirs(out_dir, exist_ok=True)
    for name, rel in TABLES.items():
        src = os.path.join(ref_root, rel)
        if not os.path.exists(src):
            print("missing", src)
            continue
        with open(src, newline="") as f:
            lines = [ln for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
        rows = list(csv.reader(lines))
        with open(os.path.join(out_dir, name), "w", newline="") as f:
            w = csv.writer(f, lineterminator="\n")
            for r in rows:
                w.writerow(r)
        print("%s: %d cases (from %s)" % (name, len(rows) - 1, rel))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
