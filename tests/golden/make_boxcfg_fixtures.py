"""Writes the configuration-file fixtures from the reference's example files (run in a container
that holds the reference; the fixtures are data excerpts, committed beside this script):

boxcfg15_w2000_he1_excerpt.cfg -- examples/NEB_Test/PrevCfg.0000 (&BOXCFG15, written by the
    reference): its 10 header lines with the atom counts set to the excerpt's, the first 30 W rows
    and the He row;
plain_w2000_excerpt.cfg -- the first 20 rows of examples/use_extern_pottable/W_2000.cfg (plain
    table: type, position in LU, velocity in cm/s, then columns the loader does not read);
cfgxyz_w2000_excerpt.cfg -- examples/Cluster_Growth_Test/W_2000.cfg (&CFGXYZ): its 5 header
    lines with &NATOM set to 20, and the first 20 rows.
"""
import os

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(REF, "NEB_Test", "PrevCfg.0000")) as fh:
    lines = fh.read().split("\n")
head = lines[:10]
rows = [l for l in lines[10:] if l.strip()]
w_rows = [l for l in rows if l.split()[0] == "1"][:30]
he_rows = [l for l in rows if l.split()[0] == "2"]
head[6] = head[6].replace("2001", str(len(w_rows) + len(he_rows)))
head[8] = head[8].replace("2000", str(len(w_rows)))
with open(os.path.join(HERE, "boxcfg15_w2000_he1_excerpt.cfg"), "w") as fh:
    fh.write("\n".join(head + w_rows + he_rows) + "\n")

with open(os.path.join(REF, "use_extern_pottable", "W_2000.cfg")) as fh:
    plain = [next(fh) for _ in range(20)]
with open(os.path.join(HERE, "plain_w2000_excerpt.cfg"), "w") as fh:
    fh.write("".join(plain))

with open(os.path.join(REF, "Cluster_Growth_Test", "W_2000.cfg")) as fh:
    xyz = [next(fh) for _ in range(25)]
xyz[1] = xyz[1].replace("2000", "20")
with open(os.path.join(HERE, "cfgxyz_w2000_excerpt.cfg"), "w") as fh:
    fh.write("".join(xyz))
