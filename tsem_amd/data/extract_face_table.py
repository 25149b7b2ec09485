"""Extract the face re-orientation table used by the global gather maps.

Reads TSEMLib's ``include/quad_info/Info_Interpolate_More`` (the table that
``TSEM::read_interp_info`` parses, src/BuildSpace.cpp:1703-1771) and writes the
rows for the face DOFs of polynomial order <= MMAX into ``face_table.txt``.
Decimal strings are copied unchanged so that parsing them yields the same
doubles the reference parses.

File layout written (one record per line, whitespace separated):

    MMAX <M> NFACE <n_face_dof(M)>
    <var> <dof> <n_e0> (<idx> <w>)*  <n_e1> (<idx> <w>)*  <n_e2> (<idx> <w>)*  <n_f> (<idx> <w>)*

``var`` is the face orientation variant 0..5, ``dof`` the face-DOF index in the
degree-blocked order ``(l1+l2-3)(l1+l2-2)/2 + l2-1``; edge terms index the DOFs
``l-2`` of the face's three edges, face terms index face DOFs.

Usage: python extract_face_table.py /root/reference/include/quad_info/Info_Interpolate_More
"""
import os
import sys

MMAX = 12


def n_face(m):
    return (m - 2) * (m - 3) // 2 + m - 2


def main(src):
    toks = open(src).read().split()
    pos = 0

    def nxt():
        nonlocal pos
        t = toks[pos]
        pos += 1
        return t

    m_file = int(nxt())
    nread = n_face(m_file)
    keep = n_face(MMAX)
    rows = []
    for var in range(6):
        edges = []
        for dof in range(nread):
            rec = []
            for _ in range(3):
                cnt = int(nxt())
                pairs = [(nxt(), nxt()) for _ in range(cnt)]
                rec.append(pairs)
            edges.append(rec)
        faces = []
        for dof in range(nread):
            cnt = int(nxt())
            faces.append([(nxt(), nxt()) for _ in range(cnt)])
        for dof in range(keep):
            for ie in range(3):
                for idx, _ in edges[dof][ie]:
                    assert int(idx) <= MMAX - 2, (var, dof, idx)
            for idx, _ in faces[dof]:
                assert int(idx) < keep, (var, dof, idx)
            parts = [str(var), str(dof)]
            for ie in range(3):
                parts.append(str(len(edges[dof][ie])))
                for idx, w in edges[dof][ie]:
                    parts += [idx, w]
            parts.append(str(len(faces[dof])))
            for idx, w in faces[dof]:
                parts += [idx, w]
            rows.append(" ".join(parts))
    assert pos == len(toks), "trailing tokens"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "face_table.txt")
    with open(out, "w") as f:
        f.write("MMAX %d NFACE %d\n" % (MMAX, keep))
        f.write("\n".join(rows) + "\n")
    print("wrote", out, len(rows), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         "/root/reference/include/quad_info/Info_Interpolate_More")
