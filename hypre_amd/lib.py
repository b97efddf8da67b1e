"""ctypes loader for libhypre_mi355x.so with the C signatures of include/*.h."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhypre_mi355x.so")

_lib = None

I = C.c_int
D = C.c_double
P = C.c_void_p
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
FP = C.POINTER(C.c_float)
PP = C.POINTER(C.c_void_p)

# name: (restype, [argtypes]); handles are passed as void*
SIGNATURES = {
    # utilities
    "HYPRE_Initialize": (I, []),
    "HYPRE_Finalize": (I, []),
    "HYPRE_GetError": (I, []),
    "HYPRE_GetErrorArg": (I, []),
    "HYPRE_ClearAllErrors": (I, []),
    "HYPRE_DescribeError": (None, [I, C.c_char_p]),
    # IJ matrix
    "HYPRE_IJMatrixCreate": (I, [I, I, I, I, I, PP]),
    "HYPRE_IJMatrixDestroy": (I, [P]),
    "HYPRE_IJMatrixInitialize": (I, [P]),
    "HYPRE_IJMatrixSetObjectType": (I, [P, I]),
    "HYPRE_IJMatrixSetValues": (I, [P, I, IP, IP, IP, DP]),
    "HYPRE_IJMatrixAddToValues": (I, [P, I, IP, IP, IP, DP]),
    "HYPRE_IJMatrixAssemble": (I, [P]),
    "HYPRE_IJMatrixGetObject": (I, [P, PP]),
    "HYPRE_IJMatrixSetDiagOffdSizes": (I, [P, IP, IP]),
    "HYPRE_IJMatrixGetValues": (I, [P, I, IP, IP, IP, DP]),
    # IJ vector
    "HYPRE_IJVectorCreate": (I, [I, I, I, PP]),
    "HYPRE_IJVectorDestroy": (I, [P]),
    "HYPRE_IJVectorInitialize": (I, [P]),
    "HYPRE_IJVectorSetObjectType": (I, [P, I]),
    "HYPRE_IJVectorSetValues": (I, [P, I, IP, DP]),
    "HYPRE_IJVectorAddToValues": (I, [P, I, IP, DP]),
    "HYPRE_IJVectorAssemble": (I, [P]),
    "HYPRE_IJVectorGetValues": (I, [P, I, IP, DP]),
    "HYPRE_IJVectorGetObject": (I, [P, PP]),
    "HYPRE_IJVectorGetLocalRange": (I, [P, IP, IP]),
    # IJ files
    "HYPRE_IJMatrixRead": (I, [C.c_char_p, I, I, PP]),
    "HYPRE_IJMatrixReadMM": (I, [C.c_char_p, I, I, PP]),
    "HYPRE_IJMatrixReadBinary": (I, [C.c_char_p, I, I, PP]),
    "HYPRE_IJMatrixPrint": (I, [P, C.c_char_p]),
    "HYPRE_IJMatrixPrintBinary": (I, [P, C.c_char_p]),
    "HYPRE_IJVectorRead": (I, [C.c_char_p, I, I, PP]),
    "HYPRE_IJVectorReadBinary": (I, [C.c_char_p, I, I, PP]),
    "HYPRE_IJVectorPrint": (I, [P, C.c_char_p]),
    "HYPRE_IJVectorPrintBinary": (I, [P, C.c_char_p]),
    # ParCSR / ParVector
    "HYPRE_ParCSRMatrixDestroy": (I, [P]),
    "HYPRE_ParCSRMatrixMatvec": (I, [D, P, P, D, P]),
    "HYPRE_ParVectorCreate": (I, [I, I, IP, PP]),
    "HYPRE_ParVectorInitialize": (I, [P]),
    "HYPRE_ParVectorDestroy": (I, [P]),
    "HYPRE_ParVectorSetRandomValues": (I, [P, I]),
    "HYPRE_ParVectorSetConstantValues": (I, [P, D]),
    "HYPRE_ParVectorScale": (I, [D, P]),
    "HYPRE_ParVectorInnerProd": (I, [P, P, DP]),
    "HYPRE_MI355X_ParVectorInnerProdNoAlloc": (I, [P, P, DP]),
    "HYPRE_ParVectorCopy": (I, [P, P]),
    "HYPRE_ParVectorAxpy": (I, [D, P, P]),
    "GenerateLaplacian": (P, [I, I, I, I, I, I, I, I, I, I, DP]),
    "GenerateLaplacian27pt": (P, [I, I, I, I, I, I, I, I, I, I, DP]),
    # BoomerAMG
    "HYPRE_BoomerAMGCreate": (I, [PP]),
    "HYPRE_BoomerAMGDestroy": (I, [P]),
    "HYPRE_BoomerAMGSetup": (I, [P, P, P, P]),
    "HYPRE_BoomerAMGSolve": (I, [P, P, P, P]),
    "HYPRE_BoomerAMGSetCoarsenType": (I, [P, I]),
    "HYPRE_BoomerAMGSetInterpType": (I, [P, I]),
    "HYPRE_BoomerAMGSetStrongThreshold": (I, [P, D]),
    "HYPRE_BoomerAMGSetMaxRowSum": (I, [P, D]),
    "HYPRE_BoomerAMGSetTruncFactor": (I, [P, D]),
    "HYPRE_BoomerAMGSetPMaxElmts": (I, [P, I]),
    "HYPRE_BoomerAMGSetRelaxType": (I, [P, I]),
    "HYPRE_BoomerAMGSetCycleRelaxType": (I, [P, I, I]),
    "HYPRE_BoomerAMGSetOuterWt": (I, [P, D]),
    "HYPRE_BoomerAMGSetNumSweeps": (I, [P, I]),
    "HYPRE_BoomerAMGSetCycleNumSweeps": (I, [P, I, I]),
    "HYPRE_BoomerAMGSetMaxIter": (I, [P, I]),
    "HYPRE_BoomerAMGSetTol": (I, [P, D]),
    "HYPRE_BoomerAMGSetMaxLevels": (I, [P, I]),
    "HYPRE_BoomerAMGSetMaxCoarseSize": (I, [P, I]),
    "HYPRE_BoomerAMGSetRelaxWt": (I, [P, D]),
    "HYPRE_BoomerAMGSetPrintLevel": (I, [P, I]),
    "HYPRE_BoomerAMGSetCycleType": (I, [P, I]),
    "HYPRE_BoomerAMGSetAggNumLevels": (I, [P, I]),
    "HYPRE_BoomerAMGSetAggInterpType": (I, [P, I]),
    "HYPRE_BoomerAMGSetAggTruncFactor": (I, [P, D]),
    "HYPRE_BoomerAMGSetAggPMaxElmts": (I, [P, I]),
    "HYPRE_BoomerAMGSetNumPaths": (I, [P, I]),
    "HYPRE_BoomerAMGGetNumIterations": (I, [P, IP]),
    "HYPRE_BoomerAMGGetFinalRelativeResidualNorm": (I, [P, DP]),
    "HYPRE_MI355X_BoomerAMGGetNumLevels": (I, [P, IP]),
    # PCG
    "HYPRE_ParCSRPCGCreate": (I, [I, PP]),
    "HYPRE_ParCSRPCGDestroy": (I, [P]),
    "HYPRE_ParCSRPCGSetup": (I, [P, P, P, P]),
    "HYPRE_ParCSRPCGSolve": (I, [P, P, P, P]),
    "HYPRE_ParCSRPCGSetTol": (I, [P, D]),
    "HYPRE_ParCSRPCGSetMaxIter": (I, [P, I]),
    "HYPRE_ParCSRPCGSetTwoNorm": (I, [P, I]),
    "HYPRE_ParCSRPCGSetPrintLevel": (I, [P, I]),
    "HYPRE_ParCSRPCGSetPrecond": (I, [P, P, P, P]),
    "HYPRE_ParCSRPCGGetNumIterations": (I, [P, IP]),
    "HYPRE_ParCSRPCGGetFinalRelativeResidualNorm": (I, [P, DP]),
    # extensions
    "HYPRE_MI355X_BoomerAMGSetCyclePrecision": (I, [P, I]),
    "HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold": (I, [P, I, D]),
    "HYPRE_MI355X_BoomerAMGSetLevelRelaxWt": (I, [P, I, D]),
    "HYPRE_MI355X_BoomerAMGSetHybridPartitions": (I, [P, I, I]),
    "HYPRE_MI355X_BoomerAMGSetSolveFormat": (I, [P, I]),
    "HYPRE_MI355X_BoomerAMGSetLevelNumSweeps": (I, [P, I, I]),
    "HYPRE_MI355X_BoomerAMGClearLevelParams": (I, [P]),
    "HYPRE_MI355X_BoomerAMGSetKeepSetupData": (I, [P, I]),
    "HYPRE_MI355X_BoomerAMGGetLevelMatrix": (I, [P, I, C.c_char, IP, IP, IP, IP, IP, DP]),
    "HYPRE_MI355X_BoomerAMGGetLevelCF": (I, [P, I, IP]),
    "HYPRE_MI355X_BoomerAMGGetLevelL1": (I, [P, I, DP]),
    "HYPRE_MI355X_BoomerAMGGetLevelMeasure": (I, [P, I, DP]),
    "HYPRE_MI355X_BoomerAMGGetLevelPMISRounds": (I, [P, I, IP]),
    "HYPRE_MI355X_BoomerAMGCycleHost": (I, [P, DP, DP]),
    "HYPRE_MI355X_PCGGetResidualHistory": (I, [P, I, DP, IP]),
    "HYPRE_MI355X_PCGGetTimes": (I, [P, DP, DP]),
    "HYPRE_MI355X_CreateS": (I, [I, IP, IP, DP, D, D, IP, IP]),
    "HYPRE_MI355X_PMIS": (I, [I, IP, IP, DP, IP, IP]),
    "HYPRE_MI355X_HMIS": (I, [I, IP, IP, DP, IP, IP]),
    "HYPRE_MI355X_RSFirstPass": (I, [I, IP, IP, IP]),
    "HYPRE_MI355X_ExtPIInterp": (I, [I, IP, IP, DP, IP, IP, IP, D, I, IP, IP, IP, IP, DP]),
    "HYPRE_MI355X_RAP": (I, [I, I, IP, IP, DP, IP, IP, DP, IP, IP, IP, DP]),
    "HYPRE_MI355X_SpMV": (I, [I, I, IP, IP, DP, D, DP, D, DP]),
    "HYPRE_MI355X_HybridGS": (I, [I, IP, IP, DP, DP, DP, DP, I, D, D, I, I]),
    "HYPRE_MI355X_ParCSRMatrixGetDevice": (I, [P, IP, IP, PP, PP, PP]),
    "HYPRE_MI355X_ParCSRMatrixCopyToHost": (I, [P, IP, IP, DP]),
    "HYPRE_MI355X_ParVectorCopyFromHost": (I, [P, DP]),
    "HYPRE_MI355X_ParVectorCopyToHost": (I, [P, DP]),
    "HYPRE_MI355X_ParCSRMatrixFromHostCSR": (P, [I, I, IP, IP, DP]),
    "HYPRE_MI355X_GenerateVarDiffusion": (P, [I, I, I, D, D, I]),
    "HYPRE_MI355X_GenerateAniso7": (P, [I, D, D]),
    "HYPRE_MI355X_ParVectorSetRandomPM": (I, [P, I]),
    "HYPRE_MI355X_ParamModelCreate": (I, [I, I, I, I, PP]),
    "HYPRE_MI355X_ParamModelDestroy": (I, [P]),
    "HYPRE_MI355X_ParamModelNumParams": (I, [P, IP]),
    "HYPRE_MI355X_ParamModelSetParams": (I, [P, FP]),
    "HYPRE_MI355X_ParamModelGetParams": (I, [P, FP]),
    "HYPRE_MI355X_ParamModelForward": (I, [P, I, FP, FP]),
    "HYPRE_MI355X_ParamModelGradient": (I, [P, I, FP, FP, FP, FP]),
    "HYPRE_MI355X_ParamModelTrain": (I, [P, I, FP, FP, I, C.c_float, FP]),
    "HYPRE_MI355X_MatrixFeatures": (I, [P, DP]),
    "HYPRE_MI355X_BoomerAMGSetParamModel": (I, [P, P, P]),
    "HYPRE_MI355X_ParamModelSetGrid": (I, [P, I, DP, DP, IP]),
    "HYPRE_MI355X_ParamModelSetPrior": (I, [P, I, I, D]),
    "HYPRE_MI355X_SetProfilingRanges": (I, [I]),
    "HYPRE_MI355X_GetProfilingRanges": (I, [IP, IP]),
    "HYPRE_MI355X_SetKernelTiming": (I, [I]),
    "HYPRE_MI355X_SetKernelTimingClass": (I, [C.c_char_p]),
    "HYPRE_MI355X_ResetKernelTiming": (I, []),
    "HYPRE_MI355X_GetKernelTiming": (I, [C.c_char_p, DP, IP, DP]),
}


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m hypre_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def fnptr(name: str) -> int:
    """Address of an exported function (for SetPrecond)."""
    return C.cast(getattr(lib(), name), C.c_void_p).value
