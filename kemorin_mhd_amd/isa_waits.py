"""Build-time check of the split GEMM kernels' hand-counted memory waits.

The split Legendre kernels issue their global->LDS copies and the A-slice
register loads as inline assembly, outside the compiler's s_waitcnt
bookkeeping, and wait for them with explicit `s_waitcnt vmcnt(N)` counts
(vector-memory operations complete in issue order).  That is only correct if
the generated code never touches a load's destination VGPRs before a wait
that retires the load, on any path through the kernel: a register copy, a
spill, or a reuse of the register for another value would read or be
overwritten by the late data.

`check()` parses gfx950 assembly, builds each kernel's control-flow graph
(blocks end at labels and branches) and propagates the set of possible
queues of in-flight vector-memory operations to a fixed point; every
instruction that reads or writes a VGPR still owned by an in-flight load is
reported, as is any scratch (private memory) traffic in the checked kernels.

The same checks run on the split-form FFT kernels, whose LDS DMA is inline
assembly too (their other loads are the compiler's own).

The compiler's hazard recognizer does not look inside inline assembly either,
so `check()` also verifies two gfx9 wait-state rules before every vector-
memory instruction, along all paths: one wait state after an SALU write of M0
before an LDS DMA (global_load_lds_*), and five after a VALU write of an SGPR
(v_readfirstlane_b32 and the like) that the instruction reads.
"""
from __future__ import annotations

import re
import subprocess
from typing import Dict, List, Set, Tuple

VMEM = ("global_load", "global_store", "global_atomic", "scratch_load", "scratch_store",
        "buffer_load", "buffer_store", "buffer_atomic", "flat_load", "flat_store", "flat_atomic")
_REG = re.compile(r"(?<![a-z_])v\[(\d+):(\d+)\]|(?<![a-z_\[])v(\d+)\b")
_WAIT = re.compile(r"s_waitcnt\b.*\bvmcnt\((\d+)\)")
_SREG = re.compile(r"(?<![a-z_])s\[(\d+):(\d+)\]|(?<![a-z_\[])s(\d+)\b")
MAX_QUEUE = 64
M0_WAIT_STATES = 1     # SALU write of M0 -> LDS DMA
SGPR_WAIT_STATES = 5   # VALU write of an SGPR -> vector memory reading it


def _regs(text: str) -> Set[int]:
    out: Set[int] = set()
    for m in _REG.finditer(text):
        if m.group(3):
            out.add(int(m.group(3)))
        else:
            out.update(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def _sregs(text: str) -> Set[int]:
    out: Set[int] = set()
    for m in _SREG.finditer(text):
        if m.group(3):
            out.add(int(m.group(3)))
        else:
            out.update(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def _valu_sdst(s: str) -> Set[int]:
    """SGPRs a VALU instruction writes (destination, or carry-out)."""
    op = s.split()[0]
    ops = [a.strip() for a in s[len(op):].split(",")]
    out = _sregs(ops[0]) if ops else set()
    if len(ops) > 1 and ("_co_" in op or "mad_u64" in op or "mad_i64" in op or "div_scale" in op):
        out |= _sregs(ops[1])
    return out


def _wait_states(s: str) -> int:
    m = re.match(r"s_nop\s+(\d+)", s)
    return int(m.group(1)) + 1 if m else 1


def _hazards(blocks, order, succ) -> Set[str]:
    """Wait-state violations before vector-memory instructions, all paths."""
    pred: Dict[str, List[str]] = {b: [] for b in order}
    for b in order:
        for t in succ[b]:
            if t in pred:
                pred[t].append(b)
    bad: Set[str] = set()

    def walk(b: str, i: int, dist: int, sread: Set[int], lds: bool, ins: str, seen):
        # look at the instructions before index i of block b, `dist` wait
        # states away so far
        while True:
            for j in range(i - 1, -1, -1):
                s = blocks[b][j]
                op = s.split()[0]
                if lds and dist < M0_WAIT_STATES and op.startswith("s_") and \
                        re.match(r"s_\w+\s+m0\b", s):
                    bad.add(f"LDS DMA {dist} wait states after M0 write ({s}): {ins}")
                if dist < SGPR_WAIT_STATES and op.startswith("v_") and _valu_sdst(s) & sread:
                    bad.add(f"vector memory {dist} wait states after VALU SGPR write ({s}): {ins}")
                dist += _wait_states(s)
                if dist >= SGPR_WAIT_STATES:
                    return
            ps = pred[b]
            if not ps:
                return
            for p in ps[1:]:
                key = (p, dist)
                if key not in seen:
                    seen.add(key)
                    walk(p, len(blocks[p]), dist, sread, lds, ins, seen)
            b = ps[0]
            if (b, dist) in seen:
                return
            seen.add((b, dist))
            i = len(blocks[b])

    for b in order:
        for i, s in enumerate(blocks[b]):
            op = s.split()[0]
            if op.startswith(VMEM):
                walk(b, i, 0, _sregs(s[len(op):]), "lds" in op, s, set())
    return bad


def _blocks(body: str):
    blocks: Dict[str, List[str]] = {"@entry": []}
    order = ["@entry"]
    cur = "@entry"
    for raw in body.split("\n"):
        s = raw.split(";")[0].strip()
        if not s:
            continue
        lab = re.match(r"^(\.LBB\w+):", s)
        if lab:
            cur = lab.group(1)
            blocks[cur] = []
            order.append(cur)
            continue
        if s.startswith(".") or s.endswith(":"):
            continue
        blocks[cur].append(s)
        op = s.split()[0]
        if op == "s_endpgm":
            break
        if op.startswith(("s_cbranch", "s_branch")):
            cur = f"@{len(order)}"
            blocks[cur] = []
            order.append(cur)
    succ: Dict[str, List[str]] = {}
    for i, b in enumerate(order):
        nxt, fall = [], True
        for s in blocks[b]:
            op = s.split()[0]
            if op == "s_branch":
                nxt.append(s.split()[1])
                fall = False
            elif op.startswith("s_cbranch"):
                nxt.append(s.split()[1])
            elif op == "s_endpgm":
                fall = False
        if fall and i + 1 < len(order):
            nxt.append(order[i + 1])
        succ[b] = nxt
    return blocks, order, succ


def _step(q: Tuple[frozenset, ...], s: str, bad: Set[str]) -> Tuple[frozenset, ...]:
    op = s.split()[0]
    w = _WAIT.match(s)
    if w:
        n = int(w.group(1))
        return q[len(q) - n:] if len(q) > n else q
    if op.startswith("s_waitcnt"):
        return q
    args = s[len(op):]
    if op.startswith(VMEM):
        parts = [a.strip() for a in args.split(",")]
        if "load" in op and "lds" not in op:
            dst = frozenset(_regs(parts[0]))
            src = _regs(",".join(parts[1:]))
        else:
            dst, src = frozenset(), _regs(args)
        # a later load may target the registers of an earlier one still in
        # flight (loads return in issue order, so the later value lands last);
        # reading them as an address or store value may not
        if any(d & src for d in q):
            bad.add(s)
        return (q + (dst,))[-MAX_QUEUE:]
    used = _regs(args)
    if any(d & used for d in q):
        bad.add(s)
    return q


def check(asm: str, pattern: str) -> Dict[str, List[str]]:
    """Problems per kernel whose symbol contains `pattern` (empty: none)."""
    out: Dict[str, List[str]] = {}
    for body in re.split(r"\n(?=_Z\w+:)", asm):
        m = re.match(r"(_Z\w+):", body)
        if not m or pattern not in m.group(1):
            continue
        blocks, order, succ = _blocks(body)
        state: Dict[str, Set[Tuple[frozenset, ...]]] = {b: set() for b in order}
        state["@entry"].add(())
        work = ["@entry"]
        bad: Set[str] = set()
        while work:
            b = work.pop()
            for q in list(state[b]):
                for s in blocks[b]:
                    q = _step(q, s, bad)
                for t in succ[b]:
                    if t in state and q not in state[t]:
                        state[t].add(q)
                        work.append(t)
        probs = [f"uses a VGPR of an in-flight load: {s}" for s in sorted(bad)]
        probs += sorted(_hazards(blocks, order, succ))
        if "scratch_" in body:
            probs.append("scratch memory traffic (would disturb the counted waits)")
        out[m.group(1)] = probs
    return out


def check_source(hipcc: str, flags: List[str], src: str, pattern: str) -> Dict[str, List[str]]:
    """Compile `src` to gfx950 device assembly and check it."""
    dev = [f for f in flags if f != "-fPIC"]
    r = subprocess.run([hipcc, *dev, "--cuda-device-only", "-S", "-o", "-", src],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"device assembly of {src} failed:\n{r.stderr}")
    return check(r.stdout, pattern)
