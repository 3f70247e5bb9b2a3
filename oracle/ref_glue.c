/* Test-harness glue for the reference C kernels (oracle/_ref only).
 *
 * The reference sources declare the NumPy C-API table under the name
 * GPAW_ARRAY_API (c/extensions.h: PY_ARRAY_UNIQUE_SYMBOL) and expect the
 * extension module's init to fill it and to ready the Python types of the
 * objects they define.  This file does those two steps -- defines and fills
 * the table with NumPy's own import routine and calls PyType_Ready on the
 * four types -- so the reference objects (Operator, Transformer, Spline, LFC)
 * and fd_precond can be used from Python through ctypes.  It contains no
 * reference code.
 */
#include <Python.h>
#define PY_ARRAY_UNIQUE_SYMBOL GPAW_ARRAY_API
#include <numpy/arrayobject.h>

extern PyTypeObject OperatorType;
extern PyTypeObject TransformerType;
extern PyTypeObject SplineType;
extern PyTypeObject LFCType;

int gpawref_init(void)
{
    if (_import_array() < 0) {
        PyErr_Print();
        return -1;
    }
    if (PyType_Ready(&OperatorType) < 0 || PyType_Ready(&TransformerType) < 0
        || PyType_Ready(&SplineType) < 0 || PyType_Ready(&LFCType) < 0) {
        PyErr_Print();
        return -2;
    }
    return 0;
}
