// cpu_module.cpp -- the CPU device module _libepseon_cpu.
//
// Same surface as the reference's CPU module (cpp/cpu/source/libcpu.cpp:8-36):
// one function, greet(), returning a greeting string.  The reference has no
// CPU numerics for the Numerov path; the FP64 CPU restatement the tests
// compare against lives under oracle/ and is test infrastructure, not a
// device path.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

static PyObject* greet(PyObject*, PyObject*)
{
    return PyUnicode_FromString("Hello, World from C++!");
}

static PyMethodDef kMethods[] = {
    {"greet", greet, METH_NOARGS, "Greet the world."},
    {nullptr, nullptr, 0, nullptr},
};

static struct PyModuleDef kModule = {
    PyModuleDef_HEAD_INIT, "_libepseon_cpu", "CPU device module.", -1, kMethods,
    nullptr, nullptr, nullptr, nullptr,
};

PyMODINIT_FUNC PyInit__libepseon_cpu(void) { return PyModule_Create(&kModule); }
