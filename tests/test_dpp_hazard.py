"""Wait states before the hand-issued v_fmac_f64_dpp of the inverse DWT.

The vector-ALU inverse DWT issues `v_fmac_f64_dpp` through inline asm, and the
compiler's hazard recognizer does not look inside inline asm.  On gfx9 a VALU
write of a VGPR followed by a DPP read of it needs two wait states in between.
This test disassembles the gfx950 code object built from dwt.hip (code built
here) and checks every v_fmac_f64_dpp: walking back along every control-flow
path, no VALU instruction that writes a register of the DPP source (src0)
may sit fewer than two wait states before it.  A compiler that moved a copy or
reload of that register next to the asm would fail here, not silently in the
results.  CPU only.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "pfsoftlib_amd", "_build", "dwt.o")
LLVM = "/opt/rocm/llvm/bin"
WAIT_STATES = 2

INSN = re.compile(r"^\s+(\S+)\s*(.*?)\s*//\s*([0-9A-Fa-f]+):")
BRANCH = re.compile(r"^s_(c?branch\w*)$")


def regs(tok):
    """VGPR numbers named by one operand token (v7, v[4:5]); empty otherwise."""
    m = re.fullmatch(r"v(\d+)", tok)
    if m:
        return {int(m.group(1))}
    m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    return set()


def disassemble():
    with tempfile.TemporaryDirectory() as d:
        fb, co = os.path.join(d, "dwt.fatbin"), os.path.join(d, "dwt.co")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fb, OBJ, os.devnull],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", "--input=" + fb,
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co], check=True,
                       capture_output=True)
        out = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--mcpu=gfx950", co], check=True,
                             capture_output=True, text=True).stdout
    insns = []
    for line in out.splitlines():
        m = INSN.match(line)
        if m:
            ops = [t for t in re.split(r"[\s,]+", m.group(2)) if t]
            insns.append((int(m.group(3), 16), m.group(1), ops, line.strip()))
    return insns


def hazards(insns):
    """(count of v_fmac_f64_dpp, list of hazards found) for parsed instructions
    (address, opcode, operands, text)."""
    # branch sources per target address (target = next pc + 4 * simm16)
    preds = {}
    for i, (a, op, ops, _) in enumerate(insns):
        if BRANCH.match(op) and ops:
            simm = int(ops[0]) & 0xFFFF
            simm = simm - 0x10000 if simm >= 0x8000 else simm
            preds.setdefault(a + 4 + 4 * simm, []).append(i)

    def falls_through(i):
        return insns[i][1] not in ("s_branch", "s_endpgm", "s_setpc_b64")

    def writes(i):
        op, ops = insns[i][1], insns[i][2]
        if not op.startswith("v_") or op.startswith(("v_cmp", "v_readlane", "v_readfirstlane")) or not ops:
            return set()
        return regs(ops[0])

    def check(i, need, src, path):
        """Walk back from instruction i (exclusive) with `need` wait states left."""
        if need <= 0:
            return []
        bad = []
        cands = []
        if i - 1 >= 0 and falls_through(i - 1):
            cands.append(i - 1)
        cands += preds.get(insns[i][0], [])
        for j in cands:
            if writes(j) & src:
                bad.append((path, insns[j][3]))
                continue
            op = insns[j][1]
            ws = int(insns[j][2][0], 0) + 1 if op == "s_nop" else 1
            bad += check(j, need - ws, src, path)
        return bad

    n = 0
    problems = []
    for i, (a, op, ops, text) in enumerate(insns):
        if op != "v_fmac_f64_dpp":
            continue
        n += 1
        src0 = regs(ops[1])
        assert src0, text
        problems += check(i, WAIT_STATES, src0, text)
    return n, problems


def test_fmac_dpp_sources_have_two_wait_states():
    assert os.path.exists(OBJ), "build the library first (__graft_entry__.build())"
    n, problems = hazards(disassemble())
    assert n > 0, "no v_fmac_f64_dpp in the inverse DWT code object"
    assert not problems, problems[:5]


def _seq(lines):
    return [(4 * i, op, ops, op) for i, (op, ops) in enumerate(lines)]


def test_checker_finds_a_planted_hazard():
    mul = ("v_mul_f64", ["v[36:37]", "v[36:37]", "v[70:71]"])
    fma = ("v_fmac_f64_dpp", ["v[30:31]", "v[36:37]", "v[80:81]", "row_newbcast:0"])
    other = ("v_add_f64", ["v[2:3]", "v[4:5]", "v[6:7]"])
    assert hazards(_seq([mul, fma]))[1]                       # no wait state
    assert hazards(_seq([mul, ("s_nop", ["0"]), fma]))[1]     # one wait state
    assert hazards(_seq([mul, other, fma]))[1]                # one unrelated instruction
    assert not hazards(_seq([mul, ("s_nop", ["1"]), fma]))[1]  # two wait states
    assert not hazards(_seq([mul, other, other, fma]))[1]
    # through a branch: the write sits right before a branch into the FMA's block
    br = ("s_branch", ["1"])  # target = 4 + 4 + 4 = address 12 (the FMA below)
    seq = _seq([mul, br, ("s_nop", ["5"]), fma])
    assert hazards(seq)[1]
