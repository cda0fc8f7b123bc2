"""Writer for the .EAR scene format, following the exporter's packing rules
(blender/render_EAR/__init__.py:154-196, pack/packblock) so scenes can be
produced without Blender:

    int  -> b'int4' + i32        float -> b'flt4' + f32
    3 floats -> b'vec3' + 3 x flt4                 (28 bytes)
    list of 3 vec3 -> b'tri ' + 3 x vec3            (88 bytes)
    str  -> b'str ' + utf-8 + 1..4 NUL to a multiple of 4
    block -> id(4) + i32 length + payload
"""
from __future__ import annotations

import struct


class Vec(list):
    """Marks a 3-float list as a vec3."""


class Tri(list):
    """Marks a list of three vec3 as a triangle."""


def pack(d) -> bytes:
    if isinstance(d, bool):
        d = int(d)
    if isinstance(d, int):
        return b"int4" + struct.pack("<i", d)
    if isinstance(d, float):
        return b"flt4" + struct.pack("<f", d)
    if isinstance(d, Tri):
        return b"tri " + b"".join(pack(Vec(v)) for v in d)
    if isinstance(d, (Vec, tuple)) and len(d) == 3:
        return b"vec3" + b"".join(pack(float(x)) for x in d)
    if isinstance(d, str):
        e = d.encode("utf8")
        return b"str " + e + b"\x00" * (4 - len(e) % 4)
    if isinstance(d, bytes):
        return d
    raise ValueError(type(d))


def block(h: str, items) -> bytes:
    body = b"".join(pack(x) for x in items)
    return h.encode("ascii") + struct.pack("<i", len(body)) + body


class EarWriter:
    def __init__(self):
        self.parts = [b".EAR", block("VRSN", [0])]

    def material(self, name, refl, refr=None, spec=None):
        items = [name] + [float(x) for x in refl]
        if refr is not None or spec is not None:
            items += [float(x) for x in (refr if refr is not None else (0.0, 0.0, 0.0))]
        if spec is not None:
            items += [float(x) for x in spec]
        self.parts.append(block("MAT ", items))

    def settings(self, absorption=(0.01, 0.02, 0.05), drylevel=1.0, samples=10 ** 5, maxthreads=8, **extra):
        items = ["debug", 0, "absorption", Vec(absorption), "drylevel", float(drylevel), "samples", int(samples),
                 "maxthreads", int(maxthreads)]
        for k, v in extra.items():
            items += [k, Vec(v) if isinstance(v, (list, tuple)) else v]
        self.parts.append(block("SET ", items))

    def keys(self, times):
        self.parts.append(block("KEYS", [float(t) for t in times]))

    def mesh(self, material, tris):
        self.parts.append(block("MESH", [material] + [Tri([Vec(v) for v in t]) for t in tris]))

    def source(self, wav, location=None, anim=None, gain=None, offset=None, mesh=None):
        """SSRC block; mesh=(material name, triangles) makes an emitting mesh
        (the reader's `mesh` element, src/SoundFile.cpp:50-53)."""
        items = [wav]
        if mesh is not None:
            items.append(block("mesh", [mesh[0]] + [Tri([Vec(v) for v in t]) for t in mesh[1]]))
        elif anim is not None:
            items.append(block("anim", [Vec(p) for p in anim]))
        else:
            items.append(Vec(location))
        if gain is not None:
            items.append(float(gain))
            if offset is not None:
                items.append(float(offset))
        self.parts.append(block("SSRC", items))

    def mono(self, wav, location=None, anim=None):
        items = [wav, 35.0, block("anim", [Vec(p) for p in anim]) if anim is not None else Vec(location)]
        self.parts.append(block("OUT1", items))

    def stereo(self, wav, location, right_ear, head_size=0.2, head_abs=(0.1, 0.3, 0.9), anim=None,
               right_anim=None):
        items = [wav, 35.0]
        items.append(block("anim", [Vec(p) for p in anim]) if anim is not None else Vec(location))
        items.append(block("anim", [Vec(p) for p in right_anim]) if right_anim is not None else Vec(right_ear))
        items += [float(head_size), Vec(head_abs)]
        self.parts.append(block("OUT2", items))

    def raw(self, data: bytes):
        self.parts.append(data)

    def bytes(self) -> bytes:
        return b"".join(self.parts)

    def save(self, path):
        with open(path, "wb") as f:
            f.write(self.bytes())


def scene_to_ear(s: dict, path: str, *, samples=10 ** 5, absorption=(0.01, 0.02, 0.05), wav="dry.wav",
                 src_mesh=None, src_gain=None, **extra):
    """Writes a scenes.* dict as a .EAR file: one MAT + MESH block per
    material; src_mesh=(material index, triangles) makes the source an
    emitting mesh; src_gain: the source's gain field."""
    import numpy as np

    w = EarWriter()
    for i, (refl, refr, spec) in enumerate(s["materials"]):
        w.material(f"m{i}", refl, refr, spec)
    w.settings(absorption=absorption, samples=samples, **extra)
    tris = np.asarray(s["tris"], np.float64)
    for i in range(len(s["materials"])):
        sel = tris[np.asarray(s["mat"]) == i]
        if len(sel):
            w.mesh(f"m{i}", sel.tolist())
    if src_mesh is not None:
        w.source(wav, mesh=(f"m{src_mesh[0]}", src_mesh[1]), gain=src_gain)
    else:
        w.source(wav, location=s["src"], gain=src_gain)
    for r in s["recorders"]:
        if r.get("kind", 0) == 1:
            raw = tuple(1.0 - h ** 0.25 for h in r["head_abs"])  # undo the ctor's max(0,(1-h)^4)
            w.stereo("out.wav", r["pos"], r["right_ear"], r["head_size"], raw)
        else:
            w.mono("out.wav", r["pos"])
    w.save(path)
