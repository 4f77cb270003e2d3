/* _rows: the list entry points' bridge to wgfft_exec_host_rows.
 *
 * GpuFft.fft/ifft take a list of signals, like the reference's
 * &[Vec<Complex<f32>>] (src/lib.rs:433-488).  Reading each numpy row's address
 * through ctypes costs about 1.2 us per row in Python, more than copying a
 * 8 KB row, so this module takes the row addresses through the buffer
 * protocol (about 0.05 us per row), holds every row's buffer for the length of
 * the call and calls the library with the GIL released.
 *
 * exec_rows(fn, handle, rows, out, n, dir) -> (rc, bad)
 *   fn     address of wgfft_exec_host_rows
 *   handle address of the wgfft_t handle
 *   rows   list or tuple of signals
 *   out    address of a C-contiguous (len(rows), n) complex64 array
 * bad is -1 after the call, else the index of the first row that is not a
 * C-contiguous buffer of n complex64 values (numpy format "Zf"); then the
 * library is not called and rc is 0. */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int (*exec_rows_fn)(void*, const float* const*, float* const*, uint64_t, uint64_t, int);

static int is_complex64(const Py_buffer* v, unsigned long long n) {
    if (v->itemsize != 8 || (unsigned long long)v->len != n * 8 || !v->format) return 0;
    const char* f = v->format;
    if (*f == '<' || *f == '=' || *f == '@') ++f;  /* native little-endian */
    return strcmp(f, "Zf") == 0;
}

static PyObject* exec_rows(PyObject* self, PyObject* args) {
    unsigned long long fn, handle, out, n;
    int dir;
    PyObject* rows;
    (void)self;
    if (!PyArg_ParseTuple(args, "KKOKKi", &fn, &handle, &rows, &out, &n, &dir)) return NULL;
    PyObject* seq = PySequence_Fast(rows, "rows must be a list or tuple");
    if (!seq) return NULL;
    const Py_ssize_t count = PySequence_Fast_GET_SIZE(seq);
    PyObject** items = PySequence_Fast_ITEMS(seq);
    Py_buffer* views = (Py_buffer*)calloc((size_t)(count ? count : 1), sizeof(Py_buffer));
    const float** in_rows = (const float**)malloc((size_t)(count ? count : 1) * sizeof(float*));
    float** out_rows = (float**)malloc((size_t)(count ? count : 1) * sizeof(float*));
    if (!views || !in_rows || !out_rows) {
        free(views);
        free(in_rows);
        free(out_rows);
        Py_DECREF(seq);
        return PyErr_NoMemory();
    }
    Py_ssize_t held = 0, bad = -1;
    for (; held < count; ++held) {
        if (PyObject_GetBuffer(items[held], &views[held], PyBUF_C_CONTIGUOUS | PyBUF_FORMAT) != 0) {
            PyErr_Clear();
            bad = held;
            break;
        }
        if (!is_complex64(&views[held], n)) {
            PyBuffer_Release(&views[held]);
            bad = held;
            break;
        }
        in_rows[held] = (const float*)views[held].buf;
        out_rows[held] = (float*)(uintptr_t)(out + (unsigned long long)held * n * 8);
    }
    int rc = 0;
    if (bad < 0) {
        Py_BEGIN_ALLOW_THREADS
        rc = ((exec_rows_fn)(uintptr_t)fn)((void*)(uintptr_t)handle, in_rows, out_rows, n, (uint64_t)count, dir);
        Py_END_ALLOW_THREADS
    }
    for (Py_ssize_t i = 0; i < held; ++i) PyBuffer_Release(&views[i]);
    free(views);
    free(in_rows);
    free(out_rows);
    Py_DECREF(seq);
    return Py_BuildValue("(in)", rc, bad);
}

static PyMethodDef methods[] = {
    {"exec_rows", exec_rows, METH_VARARGS, "Call wgfft_exec_host_rows on a list of complex64 rows."},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_rows", NULL, -1, methods, NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__rows(void) { return PyModule_Create(&module); }
