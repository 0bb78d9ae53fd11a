"""glTF 2.0 helpers for the tests, independent of libapt's C++ reader:
write_glb makes small binary glTF files; read_glb flattens one the way the
reference's loader does (default scene, node TRS / matrix, triangle
primitives, per-vertex normals or generated ones) with numpy."""
from __future__ import annotations

import json
import struct

import numpy as np

CTYPE = {5126: np.float32, 5125: np.uint32, 5123: np.uint16, 5121: np.uint8}
NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_matrix(n):
    if "matrix" in n:
        return np.array(n["matrix"], np.float64).reshape(4, 4).T
    m = np.eye(4)
    m[:3, :3] = quat_matrix(n.get("rotation", [0, 0, 0, 1])) @ np.diag(n.get("scale", [1, 1, 1]))
    m[:3, 3] = n.get("translation", [0, 0, 0])
    return m


def read_glb(path):
    b = open(path, "rb").read()
    assert b[:4] == b"glTF"
    jlen = struct.unpack("<I", b[12:16])[0]
    doc = json.loads(b[20:20 + jlen])
    off = 20 + jlen
    binc = b""
    if off + 8 <= len(b):
        blen = struct.unpack("<I", b[off:off + 4])[0]
        binc = b[off + 8:off + 8 + blen]

    def acc(i):
        a = doc["accessors"][i]
        bv = doc["bufferViews"][a["bufferView"]]
        dt = np.dtype(CTYPE[a["componentType"]])
        nc = NCOMP[a["type"]]
        start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
        stride = bv.get("byteStride", dt.itemsize * nc)
        raw = np.frombuffer(binc, np.uint8)
        view = np.lib.stride_tricks.as_strided(raw[start:], shape=(a["count"], nc * dt.itemsize),
                                               strides=(stride, 1))
        return np.ascontiguousarray(view).view(dt).reshape(a["count"], nc)

    pos, nrm, mids = [], [], []

    def visit(i, parent):
        n = doc["nodes"][i]
        w = parent @ node_matrix(n)
        if "mesh" in n:
            for p in doc["meshes"][n["mesh"]]["primitives"]:
                P = acc(p["attributes"]["POSITION"]).astype(np.float32).astype(np.float64)
                I = acc(p["indices"]).reshape(-1).astype(np.int64)
                if "NORMAL" in p["attributes"]:
                    N = acc(p["attributes"]["NORMAL"]).astype(np.float32).astype(np.float64)
                else:
                    N = np.zeros_like(P)
                    T = I.reshape(-1, 3)
                    f = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
                    ln = np.linalg.norm(f, axis=1, keepdims=True)
                    f = np.where(ln > 0, f / np.where(ln > 0, ln, 1), 0)
                    for j in range(3):
                        np.add.at(N, T[:, j], f)
                    ln = np.linalg.norm(N, axis=1, keepdims=True)
                    N = np.where(ln > 0, N / np.where(ln > 0, ln, 1), 0).astype(np.float32).astype(np.float64)
                wp = (w[:3, :3] @ P.T).T + w[:3, 3]
                nm = np.linalg.inv(w[:3, :3]).T
                wn = (nm @ N.T).T
                wn /= np.linalg.norm(wn, axis=1, keepdims=True)
                pos.append(wp[I].reshape(-1, 9))
                nrm.append(wn[I].reshape(-1, 9))
                mids.append(np.full(len(I) // 3, p.get("material", 0)))
        for c in n.get("children", []):
            visit(c, w)

    sc = doc["scenes"][doc.get("scene", 0)]
    for r in sc["nodes"]:
        visit(r, np.eye(4))
    return (np.concatenate(pos), np.concatenate(nrm), np.concatenate(mids), doc)


def write_glb(path, nodes, meshes, materials, scene_roots, extensions=(), images=(), textures=None):
    """meshes: list of lists of primitives, each a dict with positions (v,3),
    indices (t*3,), optional normals (v,3), uvs (v,2), tangents (v,4),
    material, index_type (5125 or 5123), interleaved (positions and normals
    in one strided view).  images: (bytes, mime type) pairs stored in buffer
    views; textures: list of image indices (default: one texture per image)."""
    binc = bytearray()
    accessors, views = [], []

    def add_view(data: bytes, stride=None):
        while len(binc) % 4:
            binc.append(0)
        v = {"buffer": 0, "byteOffset": len(binc), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        binc.extend(data)
        views.append(v)
        return len(views) - 1

    def add_acc(view, ctype, count, typ, offset=0, mn=None, mx=None):
        a = {"bufferView": view, "componentType": ctype, "count": count, "type": typ}
        if offset:
            a["byteOffset"] = offset
        if mn is not None:
            a["min"], a["max"] = mn, mx
        accessors.append(a)
        return len(accessors) - 1

    gl_meshes = []
    for prims in meshes:
        gp = []
        for p in prims:
            P = np.asarray(p["positions"], np.float32)
            attrs = {}
            if p.get("interleaved") and "normals" in p:
                N = np.asarray(p["normals"], np.float32)
                inter = np.concatenate([P, N], axis=1).astype(np.float32)
                v = add_view(inter.tobytes(), stride=24)
                attrs["POSITION"] = add_acc(v, 5126, len(P), "VEC3", 0, P.min(0).tolist(), P.max(0).tolist())
                attrs["NORMAL"] = add_acc(v, 5126, len(P), "VEC3", 12)
            else:
                attrs["POSITION"] = add_acc(add_view(P.tobytes()), 5126, len(P), "VEC3", 0, P.min(0).tolist(),
                                            P.max(0).tolist())
                if "normals" in p:
                    attrs["NORMAL"] = add_acc(add_view(np.asarray(p["normals"], np.float32).tobytes()), 5126,
                                              len(P), "VEC3")
            if "uvs" in p:
                attrs["TEXCOORD_0"] = add_acc(add_view(np.asarray(p["uvs"], np.float32).tobytes()), 5126, len(P),
                                              "VEC2")
            if "tangents" in p:
                attrs["TANGENT"] = add_acc(add_view(np.asarray(p["tangents"], np.float32).tobytes()), 5126, len(P),
                                           "VEC4")
            it = p.get("index_type", 5125)
            I = np.asarray(p["indices"], np.uint32 if it == 5125 else np.uint16)
            prim = {"attributes": attrs, "indices": add_acc(add_view(I.tobytes()), it, len(I), "SCALAR")}
            if "material" in p:
                prim["material"] = p["material"]
            if "mode" in p:
                prim["mode"] = p["mode"]
            gp.append(prim)
        gl_meshes.append({"primitives": gp})
    gl_images = [{"bufferView": add_view(bytes(data)), "mimeType": mime} for data, mime in images]
    while len(binc) % 4:
        binc.append(0)
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(scene_roots)}], "nodes": nodes,
           "meshes": gl_meshes, "materials": materials, "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": len(binc)}]}
    if gl_images:
        doc["images"] = gl_images
        doc["textures"] = [{"source": i} for i in (textures if textures is not None else range(len(gl_images)))]
    if extensions:
        doc["extensionsUsed"] = list(extensions)
    js = json.dumps(doc).encode()
    while len(js) % 4:
        js += b" "
    out = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(binc))
    out += struct.pack("<I4s", len(js), b"JSON") + js + struct.pack("<I4s", len(binc), b"BIN\0") + bytes(binc)
    open(path, "wb").write(out)


def box(center, half):
    """A closed box (12 triangles) wound outwards."""
    c = np.asarray(center, np.float64)
    h = np.asarray(half, np.float64)
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64) * h + c
    f = [(0, 1, 3), (0, 3, 2), (4, 6, 7), (4, 7, 5), (0, 4, 5), (0, 5, 1), (2, 3, 7), (2, 7, 6), (0, 2, 6),
         (0, 6, 4), (1, 5, 7), (1, 7, 3)]
    idx = np.array(f, np.uint32).reshape(-1)
    return v.astype(np.float32), idx


def demo_scene(path):
    """A small scene exercising the reader: a floor, two boxes under a node
    hierarchy with TRS and matrix transforms, an emissive quad, a glass
    panel; 16-bit and 32-bit indices, an interleaved view, generated normals;
    materials with the emissive-strength, transmission, ior, volume,
    clearcoat, sheen and specular extensions."""
    floor_v = np.array([[-3, 0, -3], [3, 0, -3], [3, 0, 3], [-3, 0, 3]], np.float32)
    floor_i = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    floor_n = np.tile([0, 1, 0], (4, 1)).astype(np.float32)
    bv, bi = box((0, 0.5, 0), (0.5, 0.5, 0.5))
    light_v = np.array([[-0.6, 2.5, -0.6], [0.6, 2.5, -0.6], [0.6, 2.5, 0.6], [-0.6, 2.5, 0.6]], np.float32)
    light_i = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    glass_v = np.array([[-1, 0, 0], [1, 0, 0], [1, 1.5, 0], [-1, 1.5, 0]], np.float32)
    glass_i = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    meshes = [
        [{"positions": floor_v, "indices": floor_i, "normals": floor_n, "material": 0, "interleaved": True}],
        [{"positions": bv, "indices": bi, "material": 1, "index_type": 5123}],
        [{"positions": light_v, "indices": light_i, "material": 2}],
        [{"positions": glass_v, "indices": glass_i, "normals": np.tile([0, 0, 1], (4, 1)), "material": 3}],
    ]
    materials = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.7, 0.65, 1], "metallicFactor": 0.0,
                                  "roughnessFactor": 0.8}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.2, 0.1, 1], "metallicFactor": 0.3,
                                  "roughnessFactor": 0.4},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 0.7, "clearcoatRoughnessFactor": 0.1},
                        "KHR_materials_sheen": {"sheenColorFactor": [0.2, 0.3, 0.4], "sheenRoughnessFactor": 0.6},
                        "KHR_materials_specular": {"specularFactor": 0.5,
                                                   "specularColorFactor": [1.0, 0.9, 0.8]}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [1, 1, 1, 1], "metallicFactor": 0.0},
         "emissiveFactor": [1.0, 0.8, 0.6],
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 12.0}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.9, 1.0, 0.95, 1], "metallicFactor": 0.0,
                                  "roughnessFactor": 0.05},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0},
                        "KHR_materials_ior": {"ior": 1.45},
                        "KHR_materials_volume": {"attenuationColor": [0.9, 0.8, 0.7],
                                                 "attenuationDistance": 2.0}}},
    ]
    a = np.radians(25.0)
    nodes = [
        {"mesh": 0},
        {"translation": [0.0, 0.0, -0.5], "rotation": [0, np.sin(a / 2), 0, np.cos(a / 2)],
         "children": [2, 3]},
        {"mesh": 1, "translation": [-1.2, 0, 0.3], "scale": [1.0, 1.6, 1.0]},
        {"mesh": 1, "matrix": [0.9, 0, 0.2, 0, 0, 0.9, 0, 0, -0.2, 0, 0.9, 0, 1.1, 0, 0.2, 1]},
        {"mesh": 2},
        {"mesh": 3, "translation": [0.0, 0.0, 1.6]},
    ]
    write_glb(path, nodes, meshes, materials, [0, 1, 4, 5],
              extensions=("KHR_materials_clearcoat", "KHR_materials_sheen", "KHR_materials_specular",
                          "KHR_materials_emissive_strength", "KHR_materials_transmission", "KHR_materials_ior",
                          "KHR_materials_volume"))


def uv_box(center, half, uv_scale=1.0):
    """A box with a vertex set per face (24 vertices, 12 triangles, wound
    outwards): positions, face normals, uvs spanning each face."""
    c = np.asarray(center, np.float64)
    h = np.asarray(half, np.float64)
    P, N, UV, I = [], [], [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sign
            a, b = [k for k in range(3) if k != axis]
            if sign < 0:
                a, b = b, a  # keep the winding outwards
            base = len(P)
            for (s, t) in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = np.zeros(3)
                p[axis] = sign
                p[a], p[b] = s, t
                P.append(c + p * h)
                N.append(n)
                UV.append(((s + 1) * 0.5 * uv_scale, (t + 1) * 0.5 * uv_scale))
            I += [base, base + 1, base + 2, base, base + 2, base + 3]
    return (np.array(P, np.float32), np.array(N, np.float32), np.array(UV, np.float32),
            np.array(I, np.uint32))


def _quad(p0, p1, p2, p3, normal, uv):
    P = np.array([p0, p1, p2, p3], np.float32)
    N = np.tile(np.asarray(normal, np.float32), (4, 1))
    return {"positions": P, "normals": N, "uvs": np.asarray(uv, np.float32),
            "indices": np.array([0, 1, 2, 0, 2, 3], np.uint32)}


def _png(img, ctype=2):
    from tests import png_util
    return (png_util.encode(np.asarray(img), ctype, 8), "image/png")


def _jpg(img, quality=90):
    """A baseline 4:2:0 JPEG of an RGB image (Pillow's libjpeg-turbo)."""
    import io

    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(np.asarray(img, np.uint8))).save(b, "JPEG", quality=quality, subsampling=2)
    return (b.getvalue(), "image/jpeg")


def textured_room(path, seed=0, jpeg=False):
    """A closed-ish room (no front wall) of textured materials, written as GLB:
    every material texture slot in use, uvs beyond [0, 1] (repeat addressing),
    non-power-of-two images, tangents generated by the reader, a normal map
    shared with a clearcoat normal map, an emission-textured ceiling light, a
    transmission-textured glass box and a textured metal box.  Lit by the
    light only (no sky).  With jpeg, the RGB images (all but the colour
    texture, whose alpha is used) are stored as JPEG instead of PNG."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:10, 0:12]
    checker = ((xx // 3 + yy // 2) % 2)[..., None]
    color = np.where(checker, [200, 60, 40, 255], [40, 150, 210, 140]).astype(np.uint8)   # alpha: luminance
    mr = np.stack([np.zeros((16, 16)), rng.integers(40, 256, (16, 16)), rng.integers(0, 256, (16, 16))], -1)
    yy, xx = np.mgrid[0:16, 0:16] + 0.5

    def normal_map(fx, fy, amp):
        nx = amp * np.sin(2 * np.pi * fx * xx / 16)
        ny = amp * np.cos(2 * np.pi * fy * yy / 16)
        n = np.stack([nx, ny, np.ones_like(nx)], -1)
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        return np.round((n * 0.5 + 0.5) * 255).astype(np.uint8)

    stripes = np.repeat(np.repeat(np.array([[255, 255, 255], [255, 170, 60], [90, 140, 255], [30, 30, 30]],
                                           np.uint8)[None], 1, 0), 2, 1)
    stripes = np.tile(stripes, (8, 1, 1))  # 8 x 8
    grey = lambda lo, hi, s: np.repeat(rng.integers(lo, hi, (s, s))[..., None], 3, -1)
    sheen_tint = rng.integers(60, 256, (6, 6, 3))
    trans = np.repeat((((np.mgrid[0:4, 0:4].sum(0)) % 2) * 255)[..., None], 3, -1)
    rgb = _jpg if jpeg else _png
    images = [_png(color, 6), rgb(mr), rgb(normal_map(1, 2, 0.6)), rgb(stripes), rgb(grey(80, 256, 8)),
              rgb(grey(0, 256, 8)), rgb(normal_map(3, 1, 0.9)), rgb(grey(50, 256, 5)), rgb(sheen_tint),
              rgb(trans)]
    W, H, D = 2.0, 3.0, 2.0
    prims = [
        dict(_quad((-W, 0, D), (W, 0, D), (W, 0, -D), (-W, 0, -D), (0, 1, 0),
                   [(-1.5, 1.5), (1.5, 1.5), (1.5, -1.5), (-1.5, -1.5)]), material=0),   # floor
        dict(_quad((-W, 0, -D), (W, 0, -D), (W, H, -D), (-W, H, -D), (0, 0, 1),
                   [(0, 0), (2, 0), (2, 1.5), (0, 1.5)]), material=1),                   # back wall
        dict(_quad((-W, 0, D), (-W, 0, -D), (-W, H, -D), (-W, H, D), (1, 0, 0),
                   [(0, 0), (1.3, 0), (1.3, 1.1), (0, 1.1)]), material=2),               # left wall
        dict(_quad((W, 0, -D), (W, 0, D), (W, H, D), (W, H, -D), (-1, 0, 0),
                   [(0.2, 0.1), (2.2, 0.1), (2.2, 1.6), (0.2, 1.6)]), material=3),       # right wall
        dict(_quad((-W, H, -D), (W, H, -D), (W, H, D), (-W, H, D), (0, -1, 0),
                   [(0, 0), (1, 0), (1, 1), (0, 1)]), material=4),                       # ceiling
        dict(_quad((-0.7, H - 0.01, -0.7), (0.7, H - 0.01, -0.7), (0.7, H - 0.01, 0.7), (-0.7, H - 0.01, 0.7),
                   (0, -1, 0), [(0, 0), (1, 0), (1, 1), (0, 1)]), material=5),           # light
    ]
    for center, half, mat in (((-0.7, 0.5, -0.6), (0.45, 0.5, 0.45), 6), ((0.8, 0.35, 0.2), (0.35, 0.35, 0.5), 7)):
        P, N, UV, I = uv_box(center, half, uv_scale=1.7)
        prims.append({"positions": P, "normals": N, "uvs": UV, "indices": I, "material": mat})
    materials = [
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicRoughnessTexture": {"index": 1},
                                  "metallicFactor": 0.5, "roughnessFactor": 0.9},
         "normalTexture": {"index": 2, "scale": 0.7}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.7, 0.7, 1], "metallicFactor": 0.0,
                                  "roughnessFactor": 0.5},
         "normalTexture": {"index": 2},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 1.0, "clearcoatTexture": {"index": 4},
                                                    "clearcoatRoughnessFactor": 0.6,
                                                    "clearcoatRoughnessTexture": {"index": 5},
                                                    "clearcoatNormalTexture": {"index": 6}}}},
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 0.0, "roughnessFactor": 0.7},
         "extensions": {"KHR_materials_sheen": {"sheenColorFactor": [1, 1, 1], "sheenColorTexture": {"index": 8},
                                                "sheenRoughnessFactor": 1.0,
                                                "sheenRoughnessTexture": {"index": 7}}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.5, 0.8, 1], "metallicFactor": 0.0,
                                  "roughnessFactor": 0.3},
         "normalTexture": {"index": 2, "scale": 0.5},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 0.8, "clearcoatNormalTexture": {"index": 2}}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.8, 0.8, 1], "metallicFactor": 0.0}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.8, 0.8, 1], "metallicFactor": 0.0},
         "emissiveFactor": [1.0, 1.0, 1.0], "emissiveTexture": {"index": 3},
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 8.0}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [1, 1, 1, 1], "metallicFactor": 0.0, "roughnessFactor": 0.2},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0, "transmissionTexture": {"index": 9}},
                        "KHR_materials_ior": {"ior": 1.5}}},
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 1.0, "roughnessFactor": 0.4,
                                  "metallicRoughnessTexture": {"index": 1}}},
    ]
    meshes = [[p] for p in prims]
    nodes = [{"mesh": k} for k in range(len(meshes))]
    write_glb(path, nodes, meshes, materials, list(range(len(nodes))), images=images,
              extensions=("KHR_materials_clearcoat", "KHR_materials_sheen", "KHR_materials_emissive_strength",
                          "KHR_materials_transmission", "KHR_materials_ior"))
