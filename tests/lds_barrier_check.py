"""Static check of the gfx950 device assembly: every s_barrier must be reached with
no LDS access outstanding.

A __syncthreads() orders LDS accesses of different waves only if each wave has
waited for its own LDS operations (s_waitcnt lgkmcnt(0)) before the barrier.
The compiler inserts that wait when the access precedes the barrier in the same
stretch of code, but it has been seen to omit it at a loop header whose barrier
follows an LDS store made at the end of the previous iteration: another wave
can then read the location after the barrier before the store lands (a race
that changes results from run to run).  This module parses the assembly of each
kernel into basic blocks, propagates "LDS access issued and not yet waited for"
along every control-flow edge to a fixed point, and reports the barriers that
can be reached in that state.
"""
import collections
import re

# LDS-unit instructions that do not read or write LDS memory
_NO_MEMORY = ("ds_bpermute", "ds_permute", "ds_swizzle", "ds_append", "ds_consume")
_FUNC = re.compile(r"^(_Z[^:\s]+):")
_LABEL = re.compile(r"^(\.LBB\S+):")
_BRANCH = re.compile(r"^s_(cbranch_\w+|branch)\s+(\.LBB\S+)")


def _functions(text):
    funcs, cur = collections.OrderedDict(), None
    for line in text.split("\n"):
        m = _FUNC.match(line)
        if m:
            cur = m.group(1)
            funcs[cur] = []
            continue
        if cur is not None:
            if line.strip().startswith(".Lfunc_end"):
                cur = None
                continue
            funcs[cur].append(line)
    return funcs


def _blocks(body):
    blocks, index = [{"name": "entry", "ins": []}], {}
    for line in body:
        s = line.split(";")[0].strip()
        m = _LABEL.match(s)
        if m:
            blocks.append({"name": m.group(1), "ins": []})
            index[m.group(1)] = len(blocks) - 1
        elif s and not s.startswith("."):
            blocks[-1]["ins"].append(s)
    succ = []
    for i, b in enumerate(blocks):
        out, falls = [], True
        for s in b["ins"]:
            m = _BRANCH.match(s)
            if m:
                out.append(index[m.group(2)])
                if m.group(1) == "branch":
                    falls = False
            if s.startswith(("s_endpgm", "s_setpc")):
                falls = False
        if falls and i + 1 < len(blocks):
            out.append(i + 1)
        succ.append(out)
    return blocks, succ


def _step(ins, pending, hits=None, name=None):
    for s in ins:
        op = s.split()[0]
        if op.startswith("ds_") and not op.startswith(_NO_MEMORY):
            pending = True
        elif op == "s_waitcnt" and "lgkmcnt(0)" in s:
            pending = False
        elif op == "s_barrier" and pending and hits is not None:
            hits.append(name)
    return pending


def unsafe_barriers(asm_text):
    """{kernel symbol: [basic blocks holding a barrier reachable with LDS pending]}"""
    found = {}
    for fn, body in _functions(asm_text).items():
        blocks, succ = _blocks(body)
        state = [None] * len(blocks)  # None: not reached yet
        state[0] = False
        changed = True
        while changed:
            changed = False
            for i, b in enumerate(blocks):
                if state[i] is None:
                    continue
                out = _step(b["ins"], state[i])
                for j in succ[i]:
                    new = out if state[j] is None else (state[j] or out)
                    if new != state[j]:
                        state[j] = new
                        changed = True
        hits = []
        for i, b in enumerate(blocks):
            if state[i] is not None:
                _step(b["ins"], state[i], hits, b["name"])
        if hits:
            found[fn] = hits
    return found
