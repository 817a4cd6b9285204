"""Known-answer tests of the WGSL front end: the shader programs and expected
outputs of the reference's shader-compiler tests
(naga-cranelift/src/tests.rs), restated as data.

Each case is a vertex entry point run for one invocation (vertex_index 0,
instance 0, no bindings unless noted); the checked value is the
@location(0) output (tests.rs:113-152, `try_exec`) or, for `expect_abort`,
the invocation's abort code (function.rs:60-68).

`eval`, `eval_bool` and `assert_wgsl` build their programs with the same
template as tests.rs:162-198.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

ABORT_KILL = 3
ABORT_OOB = 4
ABORT_DIV0 = 5
ABORT_OVERFLOW = 6


@dataclass
class Kat:
    name: str
    source: str
    out_type: str                    # "i32", "u32", "f32", "vec4i", "vec4f", "bool"
    expected: object = None          # value of @location(0); ignored when expect_abort
    expect_abort: int = 0
    stage: str = "vertex"
    vertex_count: int = 1            # invocations (vertex_index 0..n-1)
    positions: Optional[Sequence[Sequence[float]]] = None   # expected clip positions
    buffers: List[bytes] = field(default_factory=list)      # storage/uniform buffers by slot


def eval_src(expression: str, preamble: str, out_ty: str) -> str:
    """tests.rs:162-183."""
    return f"""
        struct Output {{
            @builtin(position) p: vec4f,
            @location(0) output: {out_ty},
        }}

        @vertex
        fn main(@builtin(vertex_index) vertex_index: u32) -> Output {{
            {preamble}
            return Output(vec4f(), {expression});
        }}
        """


def exec_src(body: str, out_ty: str, decls: str = "") -> str:
    return f"""
        struct Output {{
            @builtin(position) p: vec4f,
            @location(0) output: {out_ty},
        }}
        {decls}
        @vertex
        fn main() -> Output {{
            {body}
        }}
        """


def _bool(expression: str, preamble: str, expected: bool, name: str) -> Kat:
    """tests.rs:185-198: the bool is returned as u32(expression)."""
    return Kat(name, eval_src(f"u32({expression})", preamble, "u32"), "u32", 1 if expected else 0)


def kats() -> List[Kat]:
    out: List[Kat] = []
    # tests.rs:47-110 vertex_triangle
    out.append(Kat("vertex_triangle", """
    @vertex
    fn main(@builtin(vertex_index) vertex_index: u32) -> @builtin(position) vec4f {
        let x = f32(i32(vertex_index) - 1);
        let y = f32(i32(vertex_index & 1u) * 2 - 1);
        return vec4f(x, y, 0.0, 1.0);
    }
    """, "none", vertex_count=3,
        positions=[[-1.0, -1.0, 0.0, 1.0], [0.0, 1.0, 0.0, 1.0], [1.0, -1.0, 0.0, 1.0]]))
    # tests.rs:200-210
    out.append(Kat("init_variable", eval_src("a", "var a: u32 = 123;", "u32"), "u32", 123))
    out.append(Kat("store_variable", eval_src("a", "var a: u32; a = 123;", "u32"), "u32", 123))
    # tests.rs:212-235 casts
    for value, in_ty, out_ty, expected in [
            ("false", "bool", "u32", "0"), ("true", "bool", "u32", "1"), ("false", "bool", "i32", "0"),
            ("true", "bool", "i32", "1"), ("false", "bool", "f32", "0.0"), ("true", "bool", "f32", "1.0"),
            ("5", "u32", "f32", "5.0"), ("-3", "i32", "f32", "-3.0")]:
        pre = f"var input: {in_ty} = {value};\n        var output: {out_ty} = {out_ty}(input);"
        out.append(_bool(f"output == {expected}", pre, True, f"cast_{in_ty}_{value}_{out_ty}"))
    # tests.rs:237-275 binops_scalars
    for ty, left, op, right, expected in [
            ("i32", "1", "+", "1", 2), ("i32", "2", "-", "1", 1), ("i32", "1", "-", "2", -1),
            ("i32", "2", "*", "3", 6), ("i32", "2", "*", "-3", -6), ("i32", "6", "/", "2", 3),
            ("i32", "3", "/", "2", 1), ("i32", "3", "%", "2", 1),
            ("f32", "1", "+", "1", 2.0), ("f32", "2", "-", "1", 1.0), ("f32", "1", "-", "2", -1.0),
            ("f32", "2", "*", "3", 6.0), ("f32", "2", "*", "-3", -6.0), ("f32", "6", "/", "2", 3.0),
            ("f32", "3", "/", "2", 1.5), ("f32", "3", "%", "2", 1.0)]:
        pre = f"var left: {ty} = {left};\n        var right: {ty} = {right};"
        out.append(Kat(f"binop_{ty}_{left}{op}{right}", eval_src(f"left {op} right", pre, ty), ty, expected))
    # tests.rs:277-304 comparisons
    for ty, left, cmp, right, expected in [
            ("i32", "2", "==", "2", True), ("i32", "1", "==", "2", False), ("i32", "1", "!=", "2", True),
            ("i32", "1", "<", "2", True), ("i32", "2", ">", "1", True), ("i32", "1", "<=", "2", True),
            ("i32", "2", "<=", "2", True), ("i32", "2", ">=", "2", True), ("i32", "3", ">=", "2", True),
            ("i32", "-1", "<", "1", True)]:
        pre = f"var left: {ty} = {left};\n        var right: {ty} = {right};"
        out.append(_bool(f"left {cmp} right", pre, expected, f"compare_{left}{cmp}{right}"))
    # tests.rs:306-347 unops
    for ty, op, inp, expected in [("i32", "-", "123", -123), ("i32", "-", "-123", 123),
                                  ("f32", "-", "123.0", -123.0), ("f32", "-", "-123.0", 123.0),
                                  ("u32", "~", "123", (~123) & 0xFFFFFFFF)]:
        out.append(Kat(f"unop_{ty}_{op}{inp}", eval_src(f"{op} input", f"var input: {ty} = {inp};", ty), ty,
                       expected))
    for op, inp, expected in [("!", "true", False), ("!", "false", True)]:
        out.append(_bool(f"{op} input", f"var input: bool = {inp};", expected, f"unop_bool_{op}{inp}"))
    # tests.rs:349-375
    out.append(_bool("x == 1", "var x: u32; var c: bool = true; if c { x = 1; } else { x = 2; }", True, "if_true"))
    out.append(_bool("x == 2", "var x: u32; var c: bool = false; if c { x = 1; } else { x = 2; }", True, "if_false"))
    out.append(Kat("early_return", eval_src("123", "return Output(vec4f(), 456);", "u32"), "u32", 456))
    out.append(Kat("if_early_return",
                   eval_src("123", "var c: bool = true; if c { return Output(vec4f(), 456); }", "u32"), "u32", 456))
    # tests.rs:377-398
    out.append(Kat("function_call", exec_src("let output = my_sum(13, 12);\n            return Output(vec4f(), output);",
                                             "i32", "fn my_sum(a: i32, b: i32) -> i32 { return a + b; }"), "i32", 25))
    # tests.rs:438-478 constants
    out.append(Kat("scalar_constant", exec_src("var bar: i32 = foo;\n            return Output(vec4f(), bar);", "i32",
                                               "const foo: i32 = 1234 * 2;"), "i32", 2468))
    out.append(Kat("vector_constant", exec_src("var bar: vec4i = foo;\n            return Output(vec4f(), bar);",
                                               "vec4i", "const foo: vec4i = vec4i(1, 2, 3, 4);"), "vec4i", [1, 2, 3, 4]))
    # tests.rs:480-502
    out.append(Kat("loops", exec_src("""var fac: i32 = 1;
            for (var i = 2; i <= 6; i += 1) {
                fac *= i;
            }
            return Output(vec4f(), fac);""", "i32"), "i32", 720))
    # tests.rs:504-648 switch
    sw = """var a: i32 = {a};
            var b: i32;
            switch a {{
                case 1: {{ b = 2; }}
                case 2, 3: {{ b = 3; {brk} }}
                case {c4} {{ b = 4; }}
                default: {{ b = {dflt}; }}
            }}
            return Output(vec4f(), b);"""
    out.append(Kat("switch_case_single", exec_src(sw.format(a=4, brk="", c4="4:", dflt=-1), "i32"), "i32", 4))
    out.append(Kat("switch_case_multiple", exec_src(sw.format(a=5, brk="", c4="4, 5, 6:", dflt=3), "i32"), "i32", 4))
    out.append(Kat("switch_default", exec_src(sw.format(a=42, brk="", c4="4, 5, 6:", dflt=3), "i32"), "i32", 3))
    out.append(Kat("switch_break", exec_src(sw.format(a=3, brk="break; b = 123;", c4="4, 5, 6:", dflt=12), "i32"),
                   "i32", 3))
    # tests.rs:650-669
    out.append(Kat("global_variable", exec_src("foo.y = 123;\n            return Output(vec4f(), foo);", "vec4i",
                                               "var<private> foo: vec4i = vec4i(1, 2, 3, 4);"), "vec4i", [1, 123, 3, 4]))
    # tests.rs:671-758 accesses
    out.append(Kat("access_index_vector_variable_read",
                   exec_src("var foo = vec4i(1, 2, 3, 4);\n            return Output(vec4f(), foo.y);", "i32"), "i32", 2))
    out.append(Kat("access_index_vector_variable_assign",
                   exec_src("var foo: vec4i = vec4i(1, 2, 3, 4);\n            foo.y = 123;\n"
                            "            return Output(vec4f(), foo);", "vec4i"), "vec4i", [1, 123, 3, 4]))
    foo = "struct Foo { a: i32, b: i32, }"
    out.append(Kat("access_index_struct_variable_read",
                   exec_src("var foo = Foo(1, 2);\n            return Output(vec4f(), foo.b);", "i32", foo), "i32", 2))
    out.append(Kat("access_index_struct_variable_assign",
                   exec_src("var foo = Foo(1, 2);\n            foo.b = 123;\n"
                            "            return Output(vec4f(), vec4i(foo.a, foo.b, 0, 0));", "vec4i", foo),
                   "vec4i", [1, 123, 0, 0]))
    # tests.rs:792-815 division by zero aborts the invocation
    out.append(Kat("trap_divide_by_zero",
                   exec_src("var a: i32 = 123;\n            var b: i32 = 0;\n            var c: i32 = a / b;\n"
                            "            return Output(vec4f(), c);", "i32"), "i32", expect_abort=ABORT_DIV0))
    # tests.rs:900-927: discard aborts a fragment invocation with Killed
    out.append(Kat("kill", """
        @fragment
        fn main(@builtin(position) position: vec4f) -> @location(0) vec4f {
            discard;
        }
        """, "vec4f", expect_abort=ABORT_KILL, stage="fragment"))
    # tests.rs:929-973 arrays
    arr = "var<private> foo: array<i32, 4> = array<i32, 4>(4, 5, 6, 7);"
    out.append(Kat("access_global_variable_array_in_bounds_static",
                   exec_src("var out = foo[2];\n            return Output(vec4f(), out);", "i32", arr), "i32", 6))
    out.append(Kat("access_global_variable_array_in_bounds_dynamic",
                   exec_src("var out = 0;\n            for (var i = 0; i < 4; i += 1) {\n                out += foo[i];\n"
                            "            }\n            return Output(vec4f(), out);", "i32", arr), "i32", 22))
    # tests.rs:975-996 is #[ignore]d in the reference (the out-of-bounds read
    # is undefined there); this front end aborts the invocation instead
    out.append(Kat("access_global_variable_array_out_of_bounds_dynamic",
                   exec_src("var out = 0;\n            for (var i = 0; i <= 4; i += 1) {\n                out += foo[i];\n"
                            "            }\n            return Output(vec4f(), out);", "i32", arr), "i32",
                   expect_abort=ABORT_OOB))
    # tests.rs:998-1100 arrayLength of a 123-element storage array
    out.append(Kat("array_length", """
        struct Output {
            @builtin(position) p: vec4f,
            @location(0) output: u32,
        }

        @group(0)
        @binding(0)
        var<storage, read> what_len_is_this_array: array<i32>;

        @vertex
        fn main() -> Output {
            let out = arrayLength(&what_len_is_this_array);
            return Output(vec4f(), out);
        }
        """, "u32", 123, buffers=[(42).to_bytes(4, "little", signed=True) * 123]))
    # tests.rs:1102-1146 select
    for c, expected in [("true", 456), ("false", 123)]:
        out.append(Kat(f"select_{c}", exec_src(f"""var out = 0;
            var a = 123;
            var b = 456;
            var c = {c};
            out = select(a, b, c);
            return Output(vec4f(), out);""", "i32"), "i32", expected))
    # tests.rs:1148-1185
    out.append(Kat("vector_access_dynamic", exec_src("""var v: vec4i = vec4i(12, 23, 34, 45);
            var s: i32 = 0;
            for (var i = 0; i < 4; i += 1) {
                s += (i + 1) * v[i];
            }
            return Output(vec4f(), s);""", "i32"), "i32", 340))
    out.append(Kat("vector_access_static", exec_src("""var v: vec4i = vec4i(12, 23, 34, 45);
            var s: i32 = v.x + 2 * v.y + 3 * v.z + 4 * v.w;
            return Output(vec4f(), s);""", "i32"), "i32", 340))
    # tests.rs:1187-1210
    out.append(Kat("vectorized_function_argument", exec_src("""let output = do_stuff(vec4i(1, 2, 3, 4));
            return Output(vec4f(), output);""", "i32", "fn do_stuff(input: vec4i) -> i32 { return input.y; }"),
                   "i32", 2))
    # tests.rs:1212-1268 diverging returns
    out.append(Kat("return_if_else_diverging", exec_src("""let output = do_stuff(1234);
            return Output(vec4f(), output);""", "i32", """fn do_stuff(x: i32) -> i32 {
            if x == 1234 {
                return 45;
            }
            else {
                return 67;
            }
        }"""), "i32", 45))
    out.append(Kat("return_from_loop_body_diverging", exec_src("""let output = do_stuff();
            return Output(vec4f(), output);""", "i32", """fn do_stuff() -> i32 {
            loop {
                return 45;
            }
            return 123;
        }"""), "i32", 45))
    # tests.rs:1270-1377 matrix and vector products
    mat = """mat4x4f(
                vec4f(1, 1, 1, 1),
                vec4f(0, 2, 0, 2),
                vec4f(3, 2, 1, 0),
                vec4f(1, 2, 3, 4),
            )"""
    out.append(Kat("matrix_vector_product", exec_src(f"""var left = {mat};
            var right = vec4f(1, 2, 3, 4);
            let output = left * right;
            return Output(vec4f(), output);""", "vec4f"), "vec4f", [14.0, 19.0, 16.0, 21.0]))
    out.append(Kat("matrix_scalar_product", exec_src(f"""var left = {mat};
            var right: f32 = 2;
            let output = left * right;
            return Output(vec4f(), output[3]);""", "vec4f"), "vec4f", [2.0, 4.0, 6.0, 8.0]))
    out.append(Kat("vector_scalar_product", exec_src("""var left = vec4f(3, 2, 1, 0);
            var right: f32 = 2;
            let output = left * right;
            return Output(vec4f(), output);""", "vec4f"), "vec4f", [6.0, 4.0, 2.0, 0.0]))
    out.append(Kat("global_matrix_init", exec_src("""var right = vec4f(1, 2, 3, 4);
            let output = matrix * right;
            return Output(vec4f(), output);""", "vec4f", f"var<private> matrix: mat4x4f = {mat};"), "vec4f",
                   [14.0, 19.0, 16.0, 21.0]))
    return out
