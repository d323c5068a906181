"""Wait states around the learned kernels' inline-asm conversions.

The learned forward pass converts MFMA accumulators to BF16 operands with
`v_cvt_pk_bf16_f32 ... clamp`, written as inline asm (the clamp modifier is
the ReLU). The compiler's hazard recognizer does not look inside inline asm,
so the wait states between those conversions and the MFMAs come from guards
in the source, and whether they suffice depends on the schedule the compiler
picks. A schedule that packs a conversion too close to an MFMA gives wrong
network outputs without any error (a grouped variant of the learned kernel
once did: acceptance fell from 0.94 to 0.80).

This test compiles pyro_learned.hip with the library's own flags to gfx950
assembly and checks every inline-asm conversion against the rules the
compiler applies to its own instructions (read off its output for gfx950:
`s_nop 11` before a VALU access to a 32x32 MFMA result, `s_nop 1` between a
VALU write and an MFMA reading it):
  - RAW: an MFMA result read by a conversion: at least 12 states;
  - WAW: a conversion writing an in-flight MFMA's destination: at least 12;
  - WAR: a conversion writing an in-flight MFMA's SrcC: at least 15 (the
    16-pass figure, conservative for the 8-pass MFMAs used here);
  - operand: a conversion's result read by an MFMA: at least 2.
A state is one instruction, or N + 1 for `s_nop N`; branches are followed.
"""
import os
import re
import shutil
import subprocess

import pytest

from cudacode_amd import build as pyro_build

RAW_STATES = 12
WAW_STATES = 12
WAR_C_STATES = 15
OPERAND_STATES = 2
HORIZON = 24  # states: beyond the largest rule


def _regs(tok):
    m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.fullmatch(r"v(\d+)", tok)
    return {int(m.group(1))} if m else set()


class Ins:
    __slots__ = ("op", "ops", "asm_cvt", "label", "target", "states")

    def __init__(self, text):
        self.label = text[:-1] if text.endswith(":") else None
        op, _, rest = text.partition(" ")
        self.op = "" if self.label else op
        self.ops = [t.strip().split()[0] for t in rest.split(",") if t.strip()] if not self.label else []
        self.asm_cvt = op == "v_cvt_pk_bf16_f32" and " clamp" in rest
        self.target = self.ops[0] if op.startswith(("s_branch", "s_cbranch")) and self.ops else None
        if self.label:
            self.states = 0
        elif op == "s_nop":
            self.states = int(self.ops[0], 0) + 1
        else:
            self.states = 1

    def writes(self):
        return _regs(self.ops[0]) if self.ops and self.op.startswith("v_") else set()

    def reads(self):
        return set().union(*[_regs(t) for t in self.ops[1:]]) if len(self.ops) > 1 else set()

    def mfma(self):
        return self.op.startswith("v_mfma") and len(self.ops) >= 4


def parse_functions(text):
    """Function name -> list of Ins, for every kernel body in an assembly file."""
    out = {}
    for m in re.finditer(r"^(\S+):[ \t]*;[ \t]*@\S+\n(.*?)^\s*s_endpgm", text, re.S | re.M):
        lines = [ln.split(";")[0].strip() for ln in m.group(2).split("\n")]
        # keep labels (".LBB0_1:"), drop assembler directives (".p2align 6")
        out[m.group(1)] = [Ins(ln) for ln in lines if ln and (ln.endswith(":") or not ln.startswith("."))]
    return out


def _walk(body, start, budget):
    """Yield (index, states before it) along every path from `start` while
    fewer than `budget` states have passed; branches are followed."""
    labels = {ins.label: i for i, ins in enumerate(body) if ins.label}
    stack, seen = [(start, 0)], set()
    while stack:
        i, st = stack.pop()
        while i < len(body) and st < budget:
            if (i, st) in seen:
                break
            seen.add((i, st))
            ins = body[i]
            yield i, st
            st += ins.states
            if ins.op.startswith("s_endpgm") or ins.op.startswith("s_setpc"):
                break
            if ins.target is not None and ins.target in labels:
                if ins.op.startswith("s_branch"):
                    i = labels[ins.target]
                    continue
                stack.append((labels[ins.target], st))
            i += 1


def hazards(body):
    """Violations of the four rules in one function body: list of strings."""
    bad = []
    for i, ins in enumerate(body):
        if ins.mfma():
            dst = _regs(ins.ops[0])
            src_c = _regs(ins.ops[3]) - dst
            for j, st in _walk(body, i + 1, HORIZON):
                other = body[j]
                if not other.asm_cvt:
                    continue
                if st < RAW_STATES and other.reads() & dst:
                    bad.append(f"RAW {st} states: {j} reads the result of {i}")
                if st < WAW_STATES and other.writes() & dst:
                    bad.append(f"WAW {st} states: {j} writes the destination of {i}")
                if st < WAR_C_STATES and other.writes() & src_c:
                    bad.append(f"WAR {st} states: {j} writes SrcC of {i}")
        if ins.asm_cvt:
            out = ins.writes()
            for j, st in _walk(body, i + 1, OPERAND_STATES):
                other = body[j]
                if other.mfma() and set().union(*[_regs(t) for t in other.ops[1:4]]) & out:
                    bad.append(f"operand {st} states: MFMA {j} reads the conversion {i}")
    return bad


def _listing(*lines):
    return parse_functions("k:  ; @k\n" + "\n".join("\t" + ln for ln in lines) + "\n\ts_endpgm\n")["k"]


MFMA = "v_mfma_f32_32x32x16_bf16 v[34:49], v[2:5], v[6:9], v[34:49]"


@pytest.mark.parametrize("lines, kind", [
    ((MFMA, "s_nop 9", "v_cvt_pk_bf16_f32 v10, v34, v35 clamp"), "RAW"),
    ((MFMA, "v_cvt_pk_bf16_f32 v38, v58, v59 clamp"), "WAW"),
    (("v_mfma_f32_32x32x16_bf16 v[50:65], v[2:5], v[6:9], v[70:85]", "s_nop 3",
      "v_cvt_pk_bf16_f32 v70, v58, v59 clamp"), "WAR"),
    (("v_cvt_pk_bf16_f32 v9, v58, v59 clamp", "s_nop 0", MFMA), "operand"),
    # across a backward branch: the loop's MFMA, then its head's conversion
    (("s_mov_b32 s0, 4", ".LBB0_1:", "v_cvt_pk_bf16_f32 v40, v34, v35 clamp", "s_nop 12",
      MFMA, "s_sub_i32 s0, s0, 1", "s_cbranch_scc1 .LBB0_1"), "RAW"),
])
def test_checker_finds_planted_hazards(lines, kind):
    found = hazards(_listing(*lines))
    assert any(f.startswith(kind) for f in found), found


def test_checker_passes_guarded_sequences():
    ok = _listing(MFMA, "s_nop 12", "v_cvt_pk_bf16_f32 v40, v34, v35 clamp", "v_cvt_pk_bf16_f32 v41, v36, v37 clamp",
                  "v_cvt_pk_bf16_f32 v42, v38, v39 clamp", "v_cvt_pk_bf16_f32 v43, v40, v41 clamp",
                  "s_nop 1", "v_mfma_f32_32x32x16_bf16 v[50:65], v[2:5], v[40:43], v[50:65]")
    assert hazards(ok) == []
    # the compiler's own conversion (no clamp) is its business, not checked
    assert hazards(_listing(MFMA, "v_cvt_pk_bf16_f32 v38, v58, v59")) == []


def _compile_learned(tmp_path):
    hipcc = pyro_build.hipcc() if (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")) else None
    if hipcc is None:
        pytest.skip("hipcc not available")
    src = "pyro_learned.hip"
    flags = [f for f in pyro_build.HIP_FLAGS if f != "-fPIC"] + pyro_build.EXTRA_FLAGS.get(src, [])
    out = tmp_path / "learned.s"
    cmd = [hipcc, *flags, "-I", os.path.join(pyro_build.ROOT, "include"), "--cuda-device-only", "-S",
           os.path.join(pyro_build.CSRC, src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def test_learned_kernels_keep_the_wait_states(tmp_path):
    funcs = parse_functions(_compile_learned(tmp_path))
    with_cvt = {name: body for name, body in funcs.items() if any(i.asm_cvt for i in body)}
    # the learned phase (every lattice-size instance) and the network evaluation
    assert any("k_learned_phase" in n for n in with_cvt), sorted(funcs)
    assert any("k_mlp_eval" in n for n in with_cvt), sorted(funcs)
    problems = {name: hazards(body) for name, body in with_cvt.items()}
    problems = {k: v for k, v in problems.items() if v}
    assert not problems, problems
