"""``OM.txt`` pretty printer (``Language/Paraiso/OM/PrettyPrint.hs:30-106``).

``old=True`` prints the spellings of the older library that wrote the
shipped dumps (``Local``/``Global`` realms, ``Imm <<Type>>``), so those dumps
can be diffed against this front-end.
"""
from __future__ import annotations

from .. import annotation as A
from .graph import OM, Imm


def _pp_edges(symbol: str, pairs) -> str:
    if not pairs:
        return ""
    items = sorted(pairs)
    return " ".join([symbol] + ["(%d)%d" % (pos, node) for pos, node in items])


def _pp_anot(anot: list, with_alive: bool = True) -> list:
    out = []
    for a in A.to_list(A.Allocation, anot):
        out.append(a.value)
    for v in A.to_list(A.Valid, anot):
        out.append(v.pp())
    for d in A.to_list(A.Direct, anot):
        out.append("Depend.Direct [%s]" % ",".join(str(x) for x in d.nodes))
    for d in A.to_list(A.Indirect, anot):
        out.append("Depend.Indirect [%s]" % ",".join(str(x) for x in d.nodes))
    if [a for a in A.to_list(A.Allocation, anot)] == [A.Allocation.Manifest]:
        for c in A.to_list(A.Calc, anot):
            out.append("Depend.Calc [%s]" % ",".join(str(x) for x in sorted(c.nodes)))
    if with_alive:
        for al in A.to_list(A.Alive, anot):
            out.append("Alive %s" % ("True" if al.alive else "False"))
    for k in A.to_list(A.KernelWriteGroup, anot):
        out.append("KernelWriteGroup {getKernelGroupID = %d}" % k.gid)
    for k in A.to_list(A.OMWriteGroup, anot):
        out.append("OMWriteGroup {getOMGroupID = %d}" % k.gid)
    return ["  " + s for s in out]


def pretty_print(om: OM, old: bool = False, with_anot: bool = True,
                 with_alive: bool = True) -> str:
    lines = ["OM name: " + om.name, "** Static Variables"]
    for s in om.setup.static_values:
        lines.append('Named (Name "%s") (%s)' % (s.name, s.dyn.show(old)))
    lines.append("")
    lines.append("** Kernels")
    for k in om.kernels:
        block = ["*** Kernel name: " + k.name]
        g = k.graph
        for i, node in enumerate(g.nodes):
            if node.is_value:
                label = node.dyn.show(old)
            elif isinstance(node.inst, Imm):
                label = node.inst.show(old)
            else:
                label = node.inst.show()
            ins = [(pos, p) for pos, p in enumerate(g.pre(i))]
            outs = [(pos, j) for j, pos in g.succ_edges(i)]
            block.append(" ".join([str(i), label, _pp_edges("<-", ins), _pp_edges("->", outs)]))
            if with_anot:
                block.extend(_pp_anot(node.anot, with_alive))
        lines.append("\n".join(block) + "\n")
    return "\n".join(lines) + "\n"
