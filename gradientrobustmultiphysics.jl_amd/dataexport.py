"""Solution export: nodevalues and writeCSV! (src/dataexport.jl:22-77).

The node values are evaluated on the device (grmp_nodevalues: the
Identity of every block at every grid node, averaged over the cells around
the node as nodevalues! does with continuous = false); the file is written
on the host in the reference's layout:
header "X1<sep>X2<sep>" then "<name> <sep>" per block, then one line per node
with the coordinates and the block values, each as "%.6f<sep>"."""
import numpy as np

from . import _abi
from ._abi import lib


def nodevalues(block):
    """nodevalues(source::FEVectorBlock, Identity) with continuous = false
    (src/interpolations.jl:704-812): array [ncomp, nnodes]."""
    fes = block.FES
    nn = fes.xgrid.nnodes
    out = np.empty((nn, fes.FEType.ncomp))
    block.parent.ctx.check(lib().grmp_nodevalues(fes.h, block.parent.vec.h, int(block.offset), _abi.dptr(out)))
    return out.T.copy()


def writeCSV(filename, Data, names=(), separator="\t"):
    """writeCSV!(filename, Data::Array{FEVectorBlock}; names, seperator)
    (src/dataexport.jl:22-77) with the Identity operator for every block."""
    Data = list(Data)
    if not Data:
        raise _abi.GrmpError(_abi.ERR_ARG, "writeCSV: no data blocks")
    grid = Data[0].FES.xgrid
    coords, _ = grid.download()
    cols, header = [], []
    for j, blk in enumerate(Data):
        if blk.FES.xgrid is not grid:
            raise _abi.GrmpError(_abi.ERR_ARG, "writeCSV: all blocks must live on the first block's grid")
        name = names[j] if j < len(names) else f"Identity({blk.name})"
        header.append(f"{name} {separator}")
        cols.append(nodevalues(blk).T)
    vals = np.hstack([coords] + cols)
    with open(filename, "w") as io:
        io.write("".join(f"X{j + 1}{separator}" for j in range(coords.shape[1])))
        io.write("".join(header))
        io.write("\n")
        fmt = ("%.6f" + separator) * vals.shape[1]
        for row in vals:
            io.write(fmt % tuple(row))
            io.write("\n")
