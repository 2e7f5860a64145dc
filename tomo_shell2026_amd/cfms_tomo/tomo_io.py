"""ctypes helpers, string/angle utilities and optimum search of the cfms_tomo API
(mirrors cfms_tomo/tomo_io.py of the reference; behaviour kept, Linux-correct pointer widths)."""
import ctypes as ct
import datetime
import time

import numpy as np

# The reference declares element-id pointers as c_long (tomo_io.py:8), which is 32-bit on
# Windows but 64-bit on Linux; the C ABI takes int32, so int32 pointers are used here.
Cptr1i = ct.POINTER(ct.c_short)   # 16-bit pixel rows
Cptr1iL = ct.POINTER(ct.c_int32)  # element ids
Cptr1d = ct.POINTER(ct.c_float)   # float32


def np_to_Cptr1i(X):
    return X.ctypes.data_as(Cptr1i)


def np_to_Cptr1iL(X):
    return X.ctypes.data_as(Cptr1iL)


def np_to_Cptr1d(X):
    return X.ctypes.data_as(Cptr1d)


def Cptr1i_to_np(ptr, rows):
    if rows == 0:
        return []
    return np.array(ptr[:rows]).astype(np.int16)


def Cptr1iL_to_np(ptr, rows):
    if rows == 0:
        return []
    return np.array(ptr[:rows]).astype(np.int32)


def Cptr1d_to_np(ptr, rows):
    if rows == 0:
        return []
    return np.array(ptr[:rows]).astype(np.float32)


def my_FStr(float_value, precision=1):
    return "{0:.{precision}f}".format(float_value, precision=precision)


def my_toRadian(degree):
    return degree * np.pi / 180.0


def my_toDegree(radian):
    return radian * 180.0 / np.pi


FStr = np.vectorize(my_FStr)
toRadian = np.vectorize(my_toRadian)
toDegree = np.vectorize(my_toDegree)


def StartTimer(s=""):
    if len(s) > 0:
        print(s)
    return time.time()


def EndTimer(start_time, filename):
    total_time = time.time() - start_time
    print(filename + "= ", datetime.timedelta(seconds=total_time), " seconds \n")


def smallestN_indices(data1D, N, maintain_order=False):
    """Round masses to 0.1 g, then argsort ascending / descending (tomo_io.py:50-58)."""
    data1D = np.round(np.asarray(data1D) * 10.0) * 0.1
    best_IDs = data1D.argsort()
    worst_IDs = data1D.argsort()[::-1]
    if best_IDs.size >= N:
        return best_IDs[:N], worst_IDs[:N]
    return best_IDs, worst_IDs


def findOptimals(YPR, Mss, bVerbose=False, nOptimalToFind=1):
    """Best and worst orientation as TomoResult (tomo_io.py:61-80)."""
    from .tomo_result import TomoResult
    best_IDs, worst_IDs = smallestN_indices(Mss, nOptimalToFind, maintain_order=True)
    nOptimalToFind = min(best_IDs.size, nOptimalToFind)
    bests, worsts = [], []
    for i in range(nOptimalToFind):
        b_id = best_IDs[i]
        B = TomoResult([YPR[b_id, 0], YPR[b_id, 1], YPR[b_id, 2], 0, Mss[b_id], Mss[b_id]])
        w_id = worst_IDs[i]
        W = TomoResult([YPR[w_id, 0], YPR[w_id, 1], YPR[w_id, 2], 0, Mss[w_id], Mss[w_id]])
        if bVerbose:
            B.print()
            W.print()
        bests.append(B)
        worsts.append(W)
    return bests[0], worsts[0]
