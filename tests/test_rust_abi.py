"""The Rust binding (bonsai-gpu-sys/src/lib.rs) against the C headers it mirrors:
every struct's size and field offsets as the C compiler lays them out (a small
C program compiled here prints sizeof/offsetof), computed for the Rust
definitions by repr(C)'s rules; every header function declared with the same
parameter and return types; every header constant with the same value. No Rust
toolchain is needed: a drift between the two layouts (a field added to one
side only, as a missing bg_stats.train_steps once was) fails here."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = [ROOT / "include" / "bonsai_gpu.h", ROOT / "include" / "bonsai_scenes.h"]
RUST = ROOT / "bonsai-gpu-sys" / "src" / "lib.rs"


def strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_structs():
    """name -> [field names] of every typedef struct with a body."""
    out = {}
    for h in HEADERS:
        text = strip_c_comments(h.read_text())
        for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} (\w+);", text, flags=re.S):
            fields = []
            for decl in m.group(2).split(";"):
                decl = decl.strip()
                if not decl:
                    continue
                # `float a, b;` declares several fields of one type
                for k, part in enumerate(decl.split(",")):
                    fields.append(re.match(r".*?(\w+)\s*(\[\d+\])?$", part.strip()).group(1))
            out[m.group(3)] = fields
    return out


def c_layouts(tmp_path, structs):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bonsai_gpu.h"', '#include "bonsai_scenes.h"',
             "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'printf("{name} %zu %zu\\n", sizeof({name}), _Alignof({name}));')
        for f in fields:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        key, *vals = line.split()
        out[key] = tuple(int(v) for v in vals)
    return out


RUST_SCALARS = {"f32": (4, 4), "u32": (4, 4), "i32": (4, 4), "c_int": (4, 4), "u8": (1, 1), "c_uchar": (1, 1),
                "c_char": (1, 1), "u64": (8, 8), "f64": (8, 8), "usize": (8, 8), "u16": (2, 2)}


def rust_structs():
    """name -> [(field, type)] of every #[repr(C)] struct with named fields."""
    text = re.sub(r"//[^\n]*", "", RUST.read_text())
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)*pub struct (\w+) \{(.*?)\}", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(","):
            decl = decl.strip()
            if not decl:
                continue
            fm = re.match(r"(?:pub\s+)?(\w+)\s*:\s*(.+)$", decl, flags=re.S)
            fields.append((fm.group(1), " ".join(fm.group(2).split())))
        out[m.group(1)] = fields
    return out


def rust_layout(name, structs, cache):
    if name in cache:
        return cache[name][:2]
    off, align, offsets = 0, 1, {}
    for f, t in structs[name]:
        size, al = rust_type(t, structs, cache)
        off = (off + al - 1) // al * al
        offsets[f] = off
        off += size
        align = max(align, al)
    size = (off + align - 1) // align * align
    cache[name] = (size, align, offsets)
    return cache[name][:2]


def rust_type(t, structs, cache):
    t = t.strip()
    if t.startswith("*const") or t.startswith("*mut"):
        return 8, 8
    m = re.match(r"\[(.+);\s*(\d+)\]$", t)
    if m:
        size, al = rust_type(m.group(1), structs, cache)
        return size * int(m.group(2)), al
    if t in RUST_SCALARS:
        return RUST_SCALARS[t]
    if t in structs:
        return rust_layout(t, structs, cache)
    raise AssertionError(f"unknown Rust field type {t}")


def test_struct_layouts_match_the_header(tmp_path):
    cs = c_structs()
    rs = rust_structs()
    assert "bg_stats" in cs and "bg_render_desc" in cs
    missing = sorted(set(cs) - set(rs))
    assert not missing, f"structs without a Rust definition: {missing}"
    lay = c_layouts(tmp_path, cs)
    cache = {}
    for name, cfields in cs.items():
        size, align = rust_layout(name, rs, cache)
        assert (size, align) == lay[name], (name, (size, align), lay[name])
        rfields = [f for f, _ in rs[name]]
        # Rust spells C's `type` as `type_`
        assert [f.rstrip("_") for f in rfields] == cfields, (name, rfields, cfields)
        offsets = cache[name][2]
        for rf, cf in zip(rfields, cfields):
            assert offsets[rf] == lay[f"{name}.{cf}"][0], (name, cf, offsets[rf], lay[f"{name}.{cf}"])


C_SCALAR = {"int": "c_int", "uint32_t": "u32", "uint64_t": "u64", "size_t": "usize", "float": "f32",
            "double": "f64", "void": "c_void", "char": "c_char", "unsigned char": "c_uchar", "uint8_t": "u8"}


def c_type_canonical(decl):
    """A C parameter or return declaration as the Rust type it binds to."""
    decl = decl.strip()
    arr = re.search(r"\[\d*\]\s*$", decl)
    is_const = decl.startswith("const ")
    base = decl[6:] if is_const else decl
    stars = base.count("*") + (1 if arr else 0)
    base = re.sub(r"\[\d*\]\s*$", "", base).replace("*", " ").split()
    # drop the parameter name if present
    words = base
    if len(words) > 1 and words[-1] not in C_SCALAR and not words[-1].startswith("bg"):
        words = words[:-1]
    elif len(words) > 1 and words[-2] in ("unsigned",) and words[-1] == "char":
        pass
    ty = " ".join(words)
    if ty.startswith("unsigned char"):
        ty = "unsigned char"
    rust = C_SCALAR.get(ty, ty)
    if stars == 0:
        return rust
    out = rust
    for k in range(stars):
        # the outermost pointer carries the leading const of a `const T*`
        q = "*const" if (is_const and k == 0) else "*mut"
        out = f"{q} {out}"
    return out


def c_functions():
    out = {}
    for h in HEADERS:
        text = strip_c_comments(h.read_text())
        text = re.sub(r"typedef struct \w+ \{.*?\} \w+;", "", text, flags=re.S)
        for m in re.finditer(r"^\s*((?:const\s+)?[\w ]+?\**)\s*\b(bgx?_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
            ret, name, params = m.group(1).strip(), m.group(2), m.group(3)
            ps = [] if params.strip() in ("", "void") else [p for p in params.split(",")]
            out[name] = (c_type_canonical(ret), [c_type_canonical(p) for p in ps])
    return out


def rust_functions():
    text = re.sub(r"//[^\n]*", "", RUST.read_text())
    block = re.search(r'extern "C" \{(.*?)\n\}', text, flags=re.S).group(1)
    out = {}
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        name, params, ret = m.group(1), m.group(2), m.group(3)
        ps = [" ".join(p.split(":", 1)[1].split()) for p in params.split(",") if p.strip()]
        out[name] = (" ".join(ret.split()) if ret else "c_void", ps)
    return out


def test_functions_match_the_header():
    cf, rf = c_functions(), rust_functions()
    assert "bg_render_film_guided" in cf and "bgx_uniform" in cf
    missing = sorted(set(cf) - set(rf))
    assert not missing, f"header functions the Rust binding lacks: {missing}"
    for name, (ret, params) in cf.items():
        rret, rparams = rf[name]
        assert rret == ret, (name, rret, ret)
        assert rparams == params, (name, rparams, params)


def test_constants_match_the_header():
    text = "\n".join(strip_c_comments(h.read_text()) for h in HEADERS)
    rust = RUST.read_text()
    for m in re.finditer(r"#define (BG_\w+) \(?(-?\w+)u?(?: << (\d+))?u?\)?", text):
        name, val, shift = m.group(1), m.group(2).rstrip("u"), m.group(3)
        v = int(val, 0) << int(shift) if shift else int(val, 0)
        rm = re.search(rf"pub const {name}: \w+ = (-?[\w_]+)(?: << (\d+))?;", rust)
        assert rm, f"{name} missing in the Rust binding"
        rv = int(rm.group(1).replace("_", ""), 0)
        if rm.group(2):
            rv <<= int(rm.group(2))
        assert rv == v, (name, rv, v)


@pytest.mark.parametrize("path", ["bonsai-gpu-sys/build.rs", "bonsai-gpu-sys/Cargo.toml"])
def test_build_script_compiles_for_gfx950_with_the_library_flags(path):
    """build.rs drives hipcc itself (no prebuilt library), with the flags the Python build uses."""
    text = (ROOT / path).read_text()
    if path.endswith("build.rs"):
        for flag in ("gfx950", "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
                     "-packed-fp32-ops", "hipcc"):
            assert flag in text, flag
    else:
        assert 'build = "build.rs"' in text and "[dependencies]" in text
