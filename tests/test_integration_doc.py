"""CPU: the Rust binding written out in INTEGRATION.md declares every entry
point of include/randnla.h, with the same parameters in the same order and
the Rust FFI type of each C type; its #[repr(C)] structs have the header's
fields in the header's order."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C type (qualifiers and pointer stars kept, spaces normalised) -> Rust FFI type
_BASE = {"int": "c_int", "int32_t": "i32", "int64_t": "i64", "uint32_t": "u32",
         "uint64_t": "u64", "double": "f64", "float": "f32", "void": "c_void", "char": "c_char",
         "randnla_ctx": "RandnlaCtx", "randnla_rng": "RandnlaRng", "randnla_opts": "RandnlaOpts",
         "randnla_solve_info": "RandnlaSolveInfo", "randnla_ls_problem": "RandnlaLsProblem",
         "randnla_ls_info": "RandnlaLsInfo", "randnla_batch_ls": "RandnlaBatchLs"}


def rust_type(ctype):
    t = ctype.strip()
    const = t.startswith("const ")
    if const:
        t = t[len("const "):].strip()
    stars = t.count("*")
    base = t.replace("*", "").strip()
    r = _BASE[base]
    if stars == 0:
        return r
    inner = ("*const " if const else "*mut ") + r
    for _ in range(stars - 1):
        inner = "*mut " + inner
    return inner


def parse_header(text):
    """[(name, return C type, [(param C type, param name), ...])] in order."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    out = []
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(randnla_\w+)\s*\(([^()]*)\)\s*;", text):
        ret = " ".join(m.group(1).split())
        ret = ret.split(";")[-1].split("}")[-1].strip()
        params = []
        body = " ".join(m.group(3).split())
        if body and body != "void":
            for p in body.split(","):
                p = p.strip()
                arr = re.search(r"\[\s*\d*\s*\]$", p)
                if arr:                       # T name[N] is a pointer to T
                    p = p[:arr.start()].strip()
                    pm = re.match(r"(.*?)(\w+)$", p)
                    params.append((pm.group(1).strip() + " *", pm.group(2)))
                    continue
                pm = re.match(r"(.*?)(\w+)$", p)
                params.append((" ".join(pm.group(1).replace("*", " * ").split()), pm.group(2)))
        out.append((m.group(2), ret, params))
    return out


def parse_rust_block(md):
    blk = re.search(r'extern "C" \{(.*?)\n\}', md, flags=re.S).group(1)
    out = []
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)\s*(->\s*([^;]+))?;", blk, flags=re.S):
        args = " ".join(m.group(2).split())
        params = []
        if args:
            for a in args.split(","):
                a = a.strip()
                if not a:
                    continue
                n, t = a.split(":", 1)
                params.append((n.strip(), " ".join(t.split())))
        ret = " ".join(m.group(4).split()) if m.group(4) else None
        out.append((m.group(1), params, ret))
    return out


def _struct_fields(text, name):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", text)
    body = m.group(1)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        ty, names = decl.split(" ", 1)
        for nm in names.split(","):
            fields.append(re.sub(r"\[.*", "", nm.strip()))
    return fields


def _rust_struct_fields(md, name):
    m = re.search(r"pub struct " + name + r"\s*\{(.*?)\}", md, flags=re.S)
    return [f.split(":")[0].replace("pub", "").strip()
            for f in m.group(1).split(",") if ":" in f]


def test_rust_block_matches_header():
    hdr = parse_header(open(os.path.join(ROOT, "include", "randnla.h")).read())
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    rust = {name: (params, ret) for name, params, ret in parse_rust_block(md)}
    assert len(hdr) >= 60
    missing = [n for n, _, _ in hdr if n not in rust]
    assert not missing, missing
    assert set(rust) == {n for n, _, _ in hdr}, set(rust) ^ {n for n, _, _ in hdr}
    for name, ret, params in hdr:
        rparams, rret = rust[name]
        assert len(rparams) == len(params), name
        for (ct, cn), (rn, rt) in zip(params, rparams):
            assert rt == rust_type(ct), (name, cn, ct, rt)
        assert rret == (None if ret == "void" else rust_type(ret)), (name, ret, rret)


def test_rust_structs_match_header():
    text = open(os.path.join(ROOT, "include", "randnla.h")).read()
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for c, r in (("randnla_rng", "RandnlaRng"), ("randnla_opts", "RandnlaOpts"),
                 ("randnla_solve_info", "RandnlaSolveInfo"),
                 ("randnla_ls_problem", "RandnlaLsProblem"), ("randnla_ls_info", "RandnlaLsInfo")):
        assert _struct_fields(text, c) == _rust_struct_fields(md, r), c


def test_parser_pins():
    # the parser itself, on declarations of each shape the header uses
    d = parse_header("int f0(void);\nvoid randnla_a(const uint32_t ctr[4], uint32_t key,\n"
                     "  uint32_t out[2]);\nconst char* randnla_b(const randnla_ctx* c);\n"
                     "int randnla_c(randnla_batch_ls** plan, void* s);")
    assert d[0] == ("randnla_a", "void", [("const uint32_t *", "ctr"), ("uint32_t", "key"),
                                          ("uint32_t *", "out")])
    assert rust_type(d[1][1]) == "*const c_char"
    assert [rust_type(t) for t, _ in d[2][2]] == ["*mut *mut RandnlaBatchLs", "*mut c_void"]
