"""The Julia front-end (julia/EnsembleMI355X.jl) cannot run here (no Julia in
this image).  These checks hold its emitters to the contract the device code
compiles against: every C++ entry point, struct and constexpr member the
Python tracer emits for the same hooks (lane forms, isoutofdomain, Discrete /
Continuous / VectorContinuousCallback, the workgroup per-size forms and their
callbacks) appears in the Julia emitter's output templates with the same
signature, and the Julia ccall layer mirrors include/odemi.h field for field."""
import os
import re

import odemi
from odemi import problems as P
from odemi.api import trace_for
from odemi.tracer import trace_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = open(os.path.join(ROOT, "ordinarydiffeq.jl_amd", "julia", "EnsembleMI355X.jl")).read()


def _norm(s):
    return re.sub(r"\s+", " ", s)


def _signatures(src):
    """`name(args)` of every __device__ function and struct member the source
    declares, with argument lists normalised."""
    out = set()
    for m in re.finditer(r"__device__ (?:__forceinline__ |__noinline__ )?(?:bool|void|T) (\w+)\(([^)]*)\)", src):
        out.add((m.group(1), _norm(m.group(2))))
    return out


JLN = _norm(JL)


def _declared(name, args):
    """The Julia emitter writes `name` and each parameter declaration of its
    argument list (the lists are assembled from shared pieces there)."""
    if not re.search(rf"\b{name}\b", JL):
        return False
    return all(_norm(a.strip()) in JLN for a in args.split(",") if a.strip())


def _decay(du, u, p, t):
    du[0] = -p[0] * u[0]
    du[1] = u[0] - u[1]


def test_julia_lane_emitter_matches_python_forms():
    cbs = [odemi.DiscreteCallback(lambda u, t, i: t == 1.0, lambda i: None),
           odemi.ContinuousCallback(lambda u, t, i: u[0] - 0.5, lambda i: None, lambda i: None),
           odemi.VectorContinuousCallback(lambda out, u, t, i: None, lambda i, ev: None, 2)]
    src = trace_problem(_decay, 2, 1, isout=lambda u, p, t: u[0] < 0.0, callbacks=cbs).src
    sigs = _signatures(src)
    assert {"odemi_f", "odemi_jac", "odemi_tgrad", "odemi_isout", "cond", "affect", "affect_neg", "vcond",
            "vaffect"} <= {name for name, _ in sigs}
    for name, args in sigs:
        assert _declared(name, args), (name, args)
    for member in re.findall(r"static constexpr (?:int|double) ([^;]*);", src):
        for field in re.findall(r"(\w+) =", member):
            assert re.search(rf"\b{field} = \$", JL), field
    for line in ("#define ODEMI_HAS_ISOUT 1", "template <int K> struct OdemiDcb;", "template <int K> struct OdemiCcb;"):
        assert line in JL, line
    for macro in ("ODEMI_NDCB", "ODEMI_NCCB", "ODEMI_VCB_MAXLEN"):
        assert re.search(rf"#define {macro} \$", JL), macro


def _vc(out, u, t, itg):
    out[0] = u[0]
    out[1] = u[1] - 0.5


def test_julia_workgroup_emitter_matches_python_forms():
    cbs = odemi.CallbackSet(odemi.ContinuousCallback(lambda u, t, itg: u[0], None, affect_neg=lambda itg: None),
                            odemi.VectorContinuousCallback(_vc, lambda itg, ev: None, 2),
                            odemi.DiscreteCallback(lambda u, t, itg: t == 0.25, lambda itg: None))
    tr, wg = trace_for(P.CHAIN, 12, 3, algs=("rodas5p",), ns=[8, 12], isoutofdomain=lambda u, p, t: t > 5.0,
                       callback=cbs)
    assert wg
    sigs = _signatures(tr.src)
    assert {"odemi_f_wg", "odemi_jac_wg", "odemi_tgrad_wg", "odemi_jvp_wg", "odemi_isout_wg", "cond", "affect",
            "affect_neg", "vcond", "vaffect"} <= {name for name, _ in sigs}
    # every OdemiCcbWg the Python tracer writes opens with vlen (the device code reads it for
    # scalar and vector callbacks alike); so does each of the Julia emitter's two templates
    heads = re.findall(r"struct OdemiCcbWg<\d+> \{\n    static constexpr int (\w+) =", tr.src)
    assert heads == ["vlen", "vlen"], heads
    jheads = re.findall(r'struct OdemiCcbWg<\$\(k - 1\)> \{\\n    static constexpr int (\w+) =', JL)
    assert jheads == ["vlen", "vlen"], jheads
    for name, args in sigs:
        assert _declared(name, args), (name, args)
    for member in re.findall(r"static constexpr (?:int|double) ([^;]*);", tr.src):
        for field in re.findall(r"(\w+) =", member):
            assert re.search(rf"\b{field} = \$", JL), field
    for line in ("#define ODEMI_HAS_ISOUT_WG 1", "template <int K> struct OdemiDcbWg;",
                 "template <int K> struct OdemiCcbWg;", "for (int q = tid; q < $cnt; q += nt) {",
                 "const int o = $r0 + q;", "J[o * ld + o + ($d)]", "#define ODEMI_GJ_RMASK 0x$"):
        assert line in JL, line
    for macro in ("ODEMI_NDCB_WG", "ODEMI_NCCB_WG", "ODEMI_VCB_MAXLEN_WG"):
        assert re.search(rf"#define {macro} \$", JL), macro


def test_julia_args_struct_mirrors_the_c_header():
    hdr = open(os.path.join(ROOT, "include", "odemi.h")).read()
    body = hdr[hdr.index("typedef struct {"):hdr.index("} odemi_solve_args;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    cfields = re.findall(r"\b(\w+)\s*(?:,|;)", body)
    m = re.search(r"mutable struct OdemiSolveArgs\n(.*?)\n    OdemiSolveArgs\(\)", JL, re.S)
    jfields = re.findall(r"(\w+)::", m.group(1))
    assert jfields == cfields
