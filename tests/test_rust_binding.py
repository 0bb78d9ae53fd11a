"""The Rust -sys crate (rust/apt-sys/src/lib.rs) mirrors include/apt.h: every
function, struct, callback type and constant the header declares is declared
in the crate with the same name, parameter count and the corresponding types
(uint32_t -> u32, const T* -> *const T, T** -> *mut *mut T, float[3] ->
[f32; 3], ...), struct fields in the same order, and the same values.  No Rust
toolchain exists in this build, so the crate is checked as text."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "apt.h"
CRATE = ROOT / "rust" / "apt-sys" / "src" / "lib.rs"

SCALAR = {"uint32_t": "u32", "uint64_t": "u64", "float": "f32", "double": "f64", "size_t": "usize",
          "int": "c_int", "uint8_t": "u8", "char": "c_char", "void": "c_void"}


def c_type_to_rust(t: str) -> str:
    t = " ".join(t.replace("*", " * ").split())
    stars = t.count("*")
    base = t.replace("*", "").replace("const", "").split()
    assert len(base) == 1, t
    name = SCALAR.get(base[0], base[0])
    if stars == 0:
        return name
    # 'const' before the base type binds the innermost pointee
    first = "*const " if t.startswith("const") else "*mut "
    return "*mut " * (stars - 1) + first + name


def split_params(s: str):
    s = " ".join(s.split())
    if s in ("", "void"):
        return []
    return [p.strip() for p in s.split(",") if p.strip()]


def header():
    t = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    t = re.sub(r"//[^\n]*", "", t)
    funcs = {}
    for m in re.finditer(r"\n((?:const )?[\w ]+?\**)\s*(apt_\w+)\(([^;{]*?)\);", t):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3)
        ps = []
        for p in split_params(params):
            pm = re.match(r"(.+?)(\w+)$", p)
            ps.append(c_type_to_rust(pm.group(1)))
        funcs[name] = (None if ret == "void" else c_type_to_rust(ret), ps)
    structs = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \w+;", t, re.S):
        fields = []
        for f in [x.strip() for x in m.group(2).split(";") if x.strip()]:
            fm = re.match(r"(.+?)(\w+)(\[(\d+)\])?$", f)
            ty = c_type_to_rust(fm.group(1))
            if fm.group(4):
                ty = f"[{ty}; {fm.group(4)}]"
            fields.append((fm.group(2), ty))
        structs[m.group(1)] = fields
    consts = {}
    for m in re.finditer(r"#define (APT_\w+) (0x[0-9A-Fa-f]+|\d+)u?", t):
        consts[m.group(1)] = int(m.group(2).rstrip("u"), 0)
    for m in re.finditer(r"typedef enum \w+ \{(.*?)\}", t, re.S):
        for name, val in re.findall(r"(APT_\w+)\s*=\s*(\d+)", m.group(1)):
            consts[name] = int(val)
    cbs = {}
    for m in re.finditer(r"typedef void \(\*(\w+)\)\((.*?)\);", t, re.S):
        cbs[m.group(1)] = [c_type_to_rust(re.match(r"(.+?)(\w+)$", p).group(1)) for p in split_params(m.group(2))]
    return funcs, structs, consts, cbs


def crate():
    t = re.sub(r"//[^\n]*", "", CRATE.read_text())
    funcs = {}
    for m in re.finditer(r"pub fn (apt_\w+)\((.*?)\)\s*(->\s*([^;]+?))?\s*;", t, re.S):
        ps = [p.split(":", 1)[1].strip() for p in split_params(m.group(2))]
        funcs[m.group(1)] = (m.group(4).strip() if m.group(4) else None, [" ".join(p.split()) for p in ps])
    structs = {}
    for m in re.finditer(r"pub struct (\w+) \{(.*?)\}", t, re.S):
        fields = []
        for line in m.group(2).split(","):
            fm = re.match(r"\s*pub (\w+): (.+)", line.strip())
            if fm:
                fields.append((fm.group(1), " ".join(fm.group(2).split())))
        structs[m.group(1)] = fields
    consts = {}
    for name, val in re.findall(r"pub const (APT_\w+): \w+ = (0x[0-9A-Fa-f_]+|\d+);", t):
        consts[name] = int(val.replace("_", ""), 0)
    cbs = {}
    for m in re.finditer(r"pub type (apt_\w+_cb) = extern \"C\" fn\((.*?)\);", t, re.S):
        cbs[m.group(1)] = [" ".join(p.split(":", 1)[1].split()) for p in split_params(m.group(2))]
    return funcs, structs, consts, cbs


def test_every_function_is_bound_with_matching_types():
    hf, _, _, _ = header()
    rf, _, _, _ = crate()
    assert len(hf) >= 40
    missing = sorted(set(hf) - set(rf))
    assert not missing, missing
    extra = sorted(set(rf) - set(hf))
    assert not extra, extra
    for name, (ret, ps) in hf.items():
        rret, rps = rf[name]
        assert rret == ret, (name, rret, ret)
        assert rps == ps, (name, rps, ps)


def test_every_struct_is_mirrored_field_for_field():
    _, hs, _, _ = header()
    _, rs, _, _ = crate()
    assert len(hs) >= 18
    for name, fields in hs.items():
        assert name in rs, name
        assert rs[name] == fields, (name, rs[name], fields)


def test_constants_and_callbacks():
    _, _, hc, hcb = header()
    _, _, rc, rcb = crate()
    for name, val in hc.items():
        assert rc.get(name) == val, (name, rc.get(name), val)
    assert hcb == rcb


def test_checker_catches_a_wrong_binding():
    """A mismatched type in a copy of the crate text fails the comparison."""
    t = CRATE.read_text().replace("pub fn apt_release(s: *mut apt_server, id: u64)",
                                  "pub fn apt_release(s: *mut apt_server, id: u32)")
    m = re.search(r"pub fn apt_release\((.*?)\)", t, re.S)
    ps = [p.split(":", 1)[1].strip() for p in split_params(m.group(1))]
    hf, _, _, _ = header()
    assert ps != hf["apt_release"][1]
