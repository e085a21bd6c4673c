"""Inline-asm statements in the HIP kernels that issue more than one instruction and
return values in registers declare those outputs early-clobber ("=&v").

Without it the compiler may give an output the register of an input, on the
assumption that the statement reads every input before it writes any output.  A
load's result lands asynchronously while later instructions of the same statement
still read the inputs: the block pair kernel's flush read its j forces through an
address register that the first of its three loads had overwritten."""
import os
import re

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "openmm_sdm_plugin_amd", "csrc")


def _asm_statements(text):
    """(start line, body) of every `asm volatile(...)` statement, parentheses matched."""
    out = []
    for m in re.finditer(r"\basm\s+volatile\s*\(", text):
        depth, i = 1, m.end()
        while depth and i < len(text):
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        out.append((text.count("\n", 0, m.start()) + 1, text[m.end():i - 1]))
    return out


def _sections(body):
    """The asm statement's template, outputs, inputs, ... (colons inside string literals,
    such as `offset:256`, do not separate sections)."""
    parts, cur, in_str = [], "", False
    for ch in body:
        if ch == '"':
            in_str = not in_str
        if ch == ":" and not in_str:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    return parts + [cur]


def test_multi_instruction_asm_outputs_are_early_clobber():
    checked = 0
    for name in sorted(os.listdir(SRC)):
        if not name.endswith((".h", ".hip")):
            continue
        text = open(os.path.join(SRC, name)).read()
        for line, body in _asm_statements(text):
            parts = _sections(body)
            template, outputs = parts[0], parts[1] if len(parts) > 1 else ""
            lines = [piece.replace("\\t", "").strip()
                     for literal in re.findall(r'"([^"]*)"', template)
                     for piece in literal.split("\\n")]
            instructions = [x for x in lines if x and not x.startswith(";")]
            constraints = re.findall(r'"(=[^"]*)"', outputs)
            if len(instructions) > 1 and constraints:
                checked += 1
                for c in constraints:
                    assert c.startswith("=&"), f"{name}:{line}: output constraint {c!r} is not early-clobber"
    assert checked >= 1  # the pair kernel's flush is one such statement
