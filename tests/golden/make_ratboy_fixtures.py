"""Writes small excerpts of the reference's shipped asset files as fixtures
(data, not code): run once in a container that has the reference mounted.

    python tests/golden/make_ratboy_fixtures.py /root/reference/bin/Objects/HairAsset/Ratboy

  ratboy_mohawk_70.tfx       the first 70 strands of Ratboy_mohawk.tfx (header
                             rewritten for 70 strands; positions and strand UVs)
  ratboy_mohawk_70.tfxbone   the bone-name table and the first 70 strands'
                             skinning records of Ratboy_mohawk.tfxbone
  ratboy_left_hand_400.tfxmesh  Ratboy_left_hand.tfxmesh's bone block, its first
                             400 vertices and the triangles among them (renumbered)
"""
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NSTRANDS = 70
NVERTS = 400


def tfx(src):
    b = open(os.path.join(src, "Ratboy_mohawk.tfx"), "rb").read()
    version, n, V, off_pos, off_uv, off_vuv, off_thick, off_col = struct.unpack_from("<fIIIIIII", b, 0)
    header = bytearray(b[:160])
    pos = b[off_pos:off_pos + NSTRANDS * V * 16]
    uv = b[off_uv:off_uv + NSTRANDS * 8]
    struct.pack_into("<fIIIIIII", header, 0, version, NSTRANDS, V, 160, 160 + len(pos), 0, 0, 0)
    with open(os.path.join(HERE, "ratboy_mohawk_70.tfx"), "wb") as f:
        f.write(bytes(header) + pos + uv)


def tfxbone(src):
    b = open(os.path.join(src, "Ratboy_mohawk.tfxbone"), "rb").read()
    p = 0
    (nb,) = struct.unpack_from("<i", b, p)
    p += 4
    for _ in range(nb):
        _, ln = struct.unpack_from("<ii", b, p)
        p += 8 + ln
    names = b[:p]
    p += 4  # strand count
    rec = 4 + 4 * 8
    with open(os.path.join(HERE, "ratboy_mohawk_70.tfxbone"), "wb") as f:
        f.write(names + struct.pack("<i", NSTRANDS) + b[p:p + NSTRANDS * rec])


def tfxmesh(src):
    lines = open(os.path.join(src, "Ratboy_left_hand.tfxmesh")).read().splitlines()
    out, i = [], 0
    while i < len(lines):
        line = lines[i]
        tok = line.split()
        if tok and tok[0].startswith("numOfVertices"):
            nv = int(tok[1])
            out.append(f"numOfVertices {NVERTS}")
            data = [l for l in lines[i + 1:] if l and not l.startswith("#")][:nv]
            out += data[:NVERTS]
            i += 1
            k = 0
            while k < nv:
                if lines[i] and not lines[i].startswith("#"):
                    k += 1
                i += 1
            continue
        if tok and tok[0].startswith("numOfTriangles"):
            nt = int(tok[1])
            data = [l for l in lines[i + 1:] if l and not l.startswith("#")][:nt]
            tris = [t.split()[1:] for t in data]
            keep = [t for t in tris if all(int(v) < NVERTS for v in t)]
            out.append(f"numOfTriangles {len(keep)}")
            out += [f"{j} {t[0]} {t[1]} {t[2]}" for j, t in enumerate(keep)]
            break
        out.append(line)
        i += 1
    with open(os.path.join(HERE, "ratboy_left_hand_400.tfxmesh"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    src = sys.argv[1]
    tfx(src)
    tfxbone(src)
    tfxmesh(src)
    for n in sorted(os.listdir(HERE)):
        print(n, os.path.getsize(os.path.join(HERE, n)))
