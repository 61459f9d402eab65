"""Extract small golden vectors from the reference's exported force-table files.

Reads (where the reference tree is present; it never travels with the repository) its table export
  examples/use_ForceTableGen/EM_TB_WANGJUN_W-HE_2010.pair / .embd
(written by Export_ForceTable, MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:1315-1459) and keeps a
subset of rows as plain-text fixtures:

  w_ackland_pair_rows.txt  -- row index, r [A], r*V1 [eV A], -r*dV1/dr [eV], RHO1 [eV^2],
                              -dRHO1/dr [eV^2/A] of table #1 (W-W Ackland-Thetford), NTAB=10000,
                              Rmax=10 A; plus RHO5 at row 1 (the largest density of any table,
                              which fixes RHOMX = RHOSCAL * max POTB)
  w_fs_embd_rows.txt       -- row index, RHO, F1(RHO) of the .embd file (FS_TYPE: F = -sqrt(RHO))

Run: python tests/golden/make_fixtures.py  (rewrites the two fixture files next to it)
"""
import os

REF = "/root/reference/examples/use_ForceTableGen"
HERE = os.path.dirname(os.path.abspath(__file__))


def rows(path):
    out = []
    with open(path) as fh:
        for line in fh:
            s = line.split()
            if not s or s[0].startswith("&") or s[0].startswith("!"):
                continue
            out.append(s)
    return out


def main():
    pair = rows(os.path.join(REF, "EM_TB_WANGJUN_W-HE_2010.pair"))
    keep = sorted(set(list(range(0, 40)) + list(range(40, 10000, 97)) + list(range(9960, 10000))))
    with open(os.path.join(HERE, "w_ackland_pair_rows.txt"), "w") as fh:
        fh.write("# from EM_TB_WANGJUN_W-HE_2010.pair (table #1 W-W Ackland-Thetford), NTAB=10000, "
                 "Rmax=10 A\n# row r rV -rdV/dr RHO -dRHO/dr\n")
        fh.write(f"# RHO5_row1 {pair[0][20]}\n")
        for k in keep:
            r = pair[k]
            fh.write(" ".join(r[:6]) + "\n")
    embd = rows(os.path.join(REF, "EM_TB_WANGJUN_W-HE_2010.embd"))
    keep = sorted(set(list(range(0, 20)) + list(range(20, 10000, 211)) + list(range(9990, 10000))))
    with open(os.path.join(HERE, "w_fs_embd_rows.txt"), "w") as fh:
        fh.write("# from EM_TB_WANGJUN_W-HE_2010.embd (FS_TYPE), NEMBD=10000, RHOSCAL=20\n"
                 "# row RHO F1(RHO)\n")
        for k in keep:
            r = embd[k]
            fh.write(" ".join(r[:3]) + "\n")


if __name__ == "__main__":
    main()
