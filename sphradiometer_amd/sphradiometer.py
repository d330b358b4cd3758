"""The ``sphradiometer`` Python module over ctypes.

Same names as the module SWIG builds from sphradiometer.i (and as the
reference's src/python/sphradiometer.i): every C function of the public
headers, ``pick_ith_correlator_plan_fd``, and the carrays / cpointer helpers
(``new_double_array``, ``double_array_getitem`` ... ``sh_seriespp_assign``).
Objects are ctypes pointers to the structs declared in sphradiometer_amd;
their fields read like the SWIG proxies' (``plan.contents.delay_product``).
"""
import ctypes

from sphradiometer_amd import (AutoPlanFD, CDouble, NetworkPlanFD, PlanFD, ShSeries, _PROTOS, lib)

_L = lib()

# every prototyped C function under its own name
for _name in _PROTOS:
    globals()[_name] = getattr(_L, _name)


def pick_ith_correlator_plan_fd(array, i):
    """sphradiometer.i:45-47"""
    return array[i]


def _array_functions(ctype, name):
    def new(n):
        return (ctype * n)()

    def delete(a):
        return None

    def getitem(a, i):
        v = a[i]
        return complex(v.re, v.im) if isinstance(v, CDouble) else v

    def setitem(a, i, v):
        if ctype is CDouble:
            v = CDouble(complex(v).real, complex(v).imag)
        a[i] = v

    globals()["new_" + name] = new
    globals()["delete_" + name] = delete
    globals()[name + "_getitem"] = getitem
    globals()[name + "_setitem"] = setitem


_array_functions(ctypes.c_double, "double_array")
_array_functions(ctypes.POINTER(ctypes.c_double), "doublep_array")
_array_functions(CDouble, "double_complex_array")
_array_functions(ctypes.POINTER(CDouble), "double_complexp_array")


# SWIG's %pointer_functions allocate with calloc and release with free, so
# a structure made by new_*p can be handed to the library's _free functions
_libc = ctypes.CDLL(None)
_libc.calloc.restype = ctypes.c_void_p
_libc.calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


def _pointer_functions(ctype, name):
    def new():
        m = _libc.calloc(1, ctypes.sizeof(ctype))
        if not m:
            raise MemoryError("new_" + name)
        return ctypes.cast(m, ctypes.POINTER(ctype))

    def copy(v):
        p = new()
        p[0] = v if isinstance(v, ctype) or not issubclass(ctype, ctypes.Structure) else ctype(v)
        return p

    def delete(p):
        _libc.free(ctypes.cast(p, ctypes.c_void_p))

    def assign(p, v):
        p[0] = v

    def value(p):
        return p[0]

    globals()["new_" + name] = new
    globals()["copy_" + name] = copy
    globals()["delete_" + name] = delete
    globals()[name + "_assign"] = assign
    globals()[name + "_value"] = value


_pointer_functions(ctypes.c_uint, "uintp")
_pointer_functions(NetworkPlanFD, "correlator_network_plan_fdp")
_pointer_functions(AutoPlanFD, "autocorrelator_network_plan_fdp")
_pointer_functions(ShSeries, "sh_seriesp")
_pointer_functions(ctypes.POINTER(ShSeries), "sh_seriespp")
