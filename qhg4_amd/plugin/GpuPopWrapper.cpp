// The plugin exports DynPopFactory looks for in a *Wrapper.so
// (QHG4/populations/DynPopFactory.cpp:80-160; the export pattern of
// QHG4/dynpops/WrapperTemplate.cpp.tmp): getInfo() names the population
// class the library provides, createPop() builds one.  The class name is
// fixed at build time (QHG_GPU_POP_CLASS), one library per class.
#include "ArrayShare.h"
#include "GpuPop.h"
#include "IDGen.h"
#include "PopFinder.h"

#ifndef QHG_GPU_POP_CLASS
#error "build with -DQHG_GPU_POP_CLASS=\"<population class>\" (see Makefile)"
#endif

extern "C" {

const std::string getInfo() { return QHG_GPU_POP_CLASS; }

// aiSeeds is part of the template's signature but DynPopFactory's call
// passes only six arguments, so it is never read
PopBase *createPop(ArrayShare *pAS, SCellGrid *pCG, PopFinder *pPopFinder, int iLayerSize, IDGen **apIDG,
                   uint32_t *aulState, uint *) {
    ArrayShare::setInstance(pAS);
    return new GpuPop(QHG_GPU_POP_CLASS, pCG, pPopFinder, iLayerSize, apIDG, aulState);
}
}
