"""Copies the reference's Si example files that the format tests compare
against (Example-Si/input/CONTROL and output files of the run) into
tests/golden/si_files/.  These are data (the example's input and its printed
results), used by tests/test_formats.py; run from the repository root with
the reference tree at /root/reference."""
import os
import shutil

SRC = "/root/reference/Example-Si"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si_files")

FILES = ["input/CONTROL"] + ["output/" + n for n in (
    "BTE.omega", "BTE.P3", "BTE.P3_plus", "BTE.P3_minus", "BTE.P3_total", "BTE.P3_plus_total", "BTE.P3_minus_total",
    "BTE.P4", "BTE.P4_plusplus", "BTE.P4_plusmius", "BTE.P4_minusminus", "BTE.P4_total", "BTE.P4_plusplus_total",
    "BTE.P4_minusminus_total", "BTE.Numprocess_4ph", "BTE.v", "BTE.qpoints", "BTE.KappaTensorVsT_RTA",
    "BTE.KappaTensorVsT_CONV")] + ["output/T300K/" + n for n in (
        "BTE.WP3", "BTE.WP3_plus", "BTE.WP3_minus", "BTE.WP4", "BTE.WP4_plusplus", "BTE.WP4_plusminus",
        "BTE.WP4_minusminus", "BTE.w_3ph", "BTE.w_3ph_plus", "BTE.w_3ph_minus", "BTE.kappa", "BTE.kappa_tensor",
        "BTE.kappa_scalar")]


def main():
    for f in FILES:
        rel = f.split("/", 1)[1]
        rel = rel if rel.startswith("T300K/") else os.path.basename(rel)
        out = os.path.join(DST, rel)
        os.makedirs(os.path.dirname(out), exist_ok=True)
        shutil.copyfile(os.path.join(SRC, f), out)
    print(f"copied {len(FILES)} files to {DST}")


if __name__ == "__main__":
    main()
