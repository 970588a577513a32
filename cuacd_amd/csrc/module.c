/* module.c - CPython extension `cuacd_amd._gpu` over the C ABI (include/cuacd_amd.h).
 *
 * Exposes the same functions, argument order and errors as the reference's
 * `cuacd._gpu` (csrc/module.c:532-562): numpy data crosses as integer
 * addresses ("K"), one process-global context, RuntimeError("cuacd error %d:
 * %s") on failure. Built against the CPython 3.10 limited API (abi3).
 * Additional functions (no reference counterpart) expose the FP64 hull
 * volume, the device-resident entry points and the timers used by bench.py.
 */
#define PY_SSIZE_T_CLEAN
#define Py_LIMITED_API 0x030A0000
#include <Python.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "cuacd_amd.h"

static gpu_ctx_t g_ctx = NULL;

static PyObject* raise_error(int rc) {
    const char* msg = gpu_last_error(g_ctx);
    PyErr_Format(PyExc_RuntimeError, "cuacd error %d: %s", rc, msg && msg[0] ? msg : "unknown");
    return NULL;
}

#define NEED_CTX()                                                                          \
    do {                                                                                    \
        if (!g_ctx) {                                                                       \
            PyErr_SetString(PyExc_RuntimeError, "context not initialized; call init() first"); \
            return NULL;                                                                    \
        }                                                                                   \
    } while (0)

#define P(T, v) ((T)(uintptr_t)(v))

/* init(device=-1, pool_bytes=0) */
static PyObject* m_init(PyObject* self, PyObject* args, PyObject* kw) {
    static char* kwlist[] = {"device", "pool_bytes", NULL};
    int device = -1;
    unsigned long long pool = 0;
    if (!PyArg_ParseTupleAndKeywords(args, kw, "|iK", kwlist, &device, &pool)) return NULL;
    if (g_ctx) {
        gpu_destroy(g_ctx);
        g_ctx = NULL;
    }
    int rc = gpu_init(&g_ctx, device, (size_t)pool);
    if (rc) {
        PyObject* r = raise_error(rc);
        if (g_ctx) {
            gpu_destroy(g_ctx);
            g_ctx = NULL;
        }
        return r;
    }
    Py_RETURN_NONE;
}

static PyObject* g_blob = NULL;
static PyObject* m_set_fatbin(PyObject* self, PyObject* arg) {
    if (!PyBytes_Check(arg)) {
        PyErr_SetString(PyExc_TypeError, "set_fatbin() expects a bytes object");
        return NULL;
    }
    Py_INCREF(arg);
    Py_XDECREF(g_blob);
    g_blob = arg;
    gpu_set_fatbin(PyBytes_AsString(arg), (size_t)PyBytes_Size(arg));
    Py_RETURN_NONE;
}

static PyObject* m_destroy(PyObject* self, PyObject* args) {
    if (g_ctx) {
        gpu_destroy(g_ctx);
        g_ctx = NULL;
    }
    Py_RETURN_NONE;
}

static PyObject* m_heap_compact(PyObject* self, PyObject* args) {
    NEED_CTX();
    int rc = gpu_heap_compact(g_ctx);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_pool_usage(PyObject* self, PyObject* args) {
    NEED_CTX();
    return PyLong_FromSize_t(gpu_pool_usage(g_ctx));
}

static PyObject* m_heap_stats(PyObject* self, PyObject* args) {
    NEED_CTX();
    unsigned long long s[6] = {0, 0, 0, 0, 0, 0};
    if (gpu_heap_stats(g_ctx, s)) {
        PyErr_SetString(PyExc_RuntimeError, "gpu_heap_stats failed");
        return NULL;
    }
    return Py_BuildValue("(KKKKKK)", s[0], s[1], s[2], s[3], s[4], s[5]);
}

static PyObject* m_test_warp_sort(PyObject* self, PyObject* args) {
    unsigned long long pp, op;
    int total, n;
    if (!PyArg_ParseTuple(args, "KiKi", &pp, &total, &op, &n)) return NULL;
    NEED_CTX();
    int rc = gpu_test_warp_sort(g_ctx, P(int*, pp), total, P(const int*, op), n);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

/* hull_dandc(pts, total, offsets, n_hulls, max_pts, max_hv, max_ht, verts, tris, nv, nt, errors) */
static PyObject* m_hull_dandc(PyObject* self, PyObject* args) {
    unsigned long long pp, op, vp, tp, nvp, ntp, ep;
    int total, nh, mp, mhv, mht;
    if (!PyArg_ParseTuple(args, "KiKiiiiKKKKK", &pp, &total, &op, &nh, &mp, &mhv, &mht, &vp, &tp, &nvp, &ntp,
                          &ep))
        return NULL;
    NEED_CTX();
    int rc = gpu_hull_dandc(g_ctx, P(const float*, pp), total, P(const int*, op), nh, mp, mhv, mht,
                            P(float*, vp), P(int*, tp), P(int*, nvp), P(int*, ntp), P(int*, ep));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

/* hull_dandc_vol(..., errors, volumes_f64)  (no reference counterpart) */
static PyObject* m_hull_dandc_vol(PyObject* self, PyObject* args) {
    unsigned long long pp, op, vp, tp, nvp, ntp, ep, volp;
    int total, nh, mp, mhv, mht;
    if (!PyArg_ParseTuple(args, "KiKiiiiKKKKKK", &pp, &total, &op, &nh, &mp, &mhv, &mht, &vp, &tp, &nvp, &ntp,
                          &ep, &volp))
        return NULL;
    NEED_CTX();
    int rc = gpu_hull_dandc_vol(g_ctx, P(const float*, pp), total, P(const int*, op), nh, mp, mhv, mht,
                                P(float*, vp), P(int*, tp), P(int*, nvp), P(int*, ntp), P(int*, ep),
                                P(double*, volp));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_test_mesh_volume(PyObject* self, PyObject* args) {
    unsigned long long vp, tp, op;
    int nv, nt;
    if (!PyArg_ParseTuple(args, "KiKiK", &vp, &nv, &tp, &nt, &op)) return NULL;
    NEED_CTX();
    int rc = gpu_test_mesh_volume(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, P(float*, op));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_batch_mesh_volume(PyObject* self, PyObject* args) {
    unsigned long long vp, tp, top, vop, volp;
    int tv, tt, n;
    if (!PyArg_ParseTuple(args, "KiKiKKiK", &vp, &tv, &tp, &tt, &top, &vop, &n, &volp)) return NULL;
    NEED_CTX();
    int rc = gpu_batch_mesh_volume(g_ctx, P(const float*, vp), tv, P(const int*, tp), tt, P(const int*, top),
                                   vop ? P(const int*, vop) : NULL, n, P(float*, volp));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_test_plane_cut(PyObject* self, PyObject* args) {
    unsigned long long vp, tp, a, b, c, d;
    int nv, nt, ac, bc, cc, dc;
    float pa, pb, pc, pd;
    if (!PyArg_ParseTuple(args, "KiKiffffKiKiKiKi", &vp, &nv, &tp, &nt, &pa, &pb, &pc, &pd, &a, &ac, &b, &bc, &c,
                          &cc, &d, &dc))
        return NULL;
    NEED_CTX();
    int npv = 0, npt = 0, nnv = 0, nnt = 0;
    int rc = gpu_test_plane_cut(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, pa, pb, pc, pd,
                                P(float*, a), ac, P(int*, b), bc, P(float*, c), cc, P(int*, d), dc, &npv, &npt,
                                &nnv, &nnt);
    if (rc) {
        PyErr_Format(PyExc_RuntimeError, "plane_cut failed: %s", gpu_last_error(g_ctx));
        return NULL;
    }
    return Py_BuildValue("iiii", npv, npt, nnv, nnt);
}

static PyObject* m_plane_cut_batch(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, mo, pl, ovp, ovn, osz, ost;
    int nm, nc;
    if (!PyArg_ParseTuple(args, "KKKKiKKiKKKK", &v, &vo, &t, &to, &nm, &mo, &pl, &nc, &ovp, &ovn, &osz, &ost))
        return NULL;
    NEED_CTX();
    int rc = gpu_plane_cut_batch(g_ctx, P(const float*, v), P(const int*, vo), P(const int*, t), P(const int*, to),
                                 nm, P(const int*, mo), P(const float*, pl), nc, P(double*, ovp), P(double*, ovn),
                                 P(int*, osz), P(int*, ost));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_score_cuts(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, mo, pl, osc, odt, ost;
    int nm, nc;
    float tau;
    if (!PyArg_ParseTuple(args, "KKKKiKKifKKK", &v, &vo, &t, &to, &nm, &mo, &pl, &nc, &tau, &osc, &odt, &ost))
        return NULL;
    NEED_CTX();
    int rc = gpu_score_cuts(g_ctx, P(const float*, v), P(const int*, vo), P(const int*, t), P(const int*, to), nm,
                            P(const int*, mo), P(const float*, pl), nc, tau, P(float*, osc), P(double*, odt),
                            P(double*, ost));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_score_plan_create(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, mo, pl;
    int nm, nc;
    float tau;
    if (!PyArg_ParseTuple(args, "KKKKiKKif", &v, &vo, &t, &to, &nm, &mo, &pl, &nc, &tau)) return NULL;
    NEED_CTX();
    score_plan_t plan = NULL;
    int rc = gpu_score_plan_create(g_ctx, P(const float*, v), P(const int*, vo), P(const int*, t), P(const int*, to),
                                   nm, P(const int*, mo), P(const float*, pl), nc, tau, &plan);
    if (rc) return raise_error(rc);
    return PyLong_FromUnsignedLongLong((unsigned long long)(uintptr_t)plan);
}
static PyObject* m_score_plan_run(PyObject* self, PyObject* args) {
    unsigned long long p, osc, odt;
    if (!PyArg_ParseTuple(args, "KKK", &p, &osc, &odt)) return NULL;
    NEED_CTX();
    int rc = gpu_score_plan_run(g_ctx, P(score_plan_t, p), P(float*, osc), P(double*, odt));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_score_plan_destroy(PyObject* self, PyObject* args) {
    unsigned long long p;
    if (!PyArg_ParseTuple(args, "K", &p)) return NULL;
    gpu_score_plan_destroy(g_ctx, P(score_plan_t, p));
    Py_RETURN_NONE;
}

static PyObject* m_set_tuning(PyObject* self, PyObject* args) {
    int key, value;
    if (!PyArg_ParseTuple(args, "ii", &key, &value)) return NULL;
    int rc = gpu_set_tuning(g_ctx, key, value);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_get_tuning(PyObject* self, PyObject* args) {
    int key, value = 0;
    if (!PyArg_ParseTuple(args, "i", &key)) return NULL;
    int rc = gpu_get_tuning(g_ctx, key, &value);
    if (rc) return raise_error(rc);
    return PyLong_FromLong(value);
}

static PyObject* m_mlp_n_params(PyObject* self, PyObject* args) { return PyLong_FromLong(gpu_mlp_n_params()); }

static PyObject* m_mlp_create(PyObject* self, PyObject* args) {
    unsigned long long params;
    int seed;
    if (!PyArg_ParseTuple(args, "Ki", &params, &seed)) return NULL;
    NEED_CTX();
    gpu_mlp_t m = NULL;
    int rc = gpu_mlp_create(g_ctx, params ? P(const float*, params) : NULL, seed, &m);
    if (rc) return raise_error(rc);
    return PyLong_FromUnsignedLongLong((unsigned long long)(uintptr_t)m);
}

static PyObject* m_mlp_destroy(PyObject* self, PyObject* args) {
    unsigned long long h;
    if (!PyArg_ParseTuple(args, "K", &h)) return NULL;
    if (g_ctx && h) gpu_mlp_destroy(g_ctx, (gpu_mlp_t)(uintptr_t)h);
    Py_RETURN_NONE;
}

static PyObject* m_mlp_get_params(PyObject* self, PyObject* args) {
    unsigned long long h, out;
    if (!PyArg_ParseTuple(args, "KK", &h, &out)) return NULL;
    NEED_CTX();
    int rc = gpu_mlp_get_params(g_ctx, (gpu_mlp_t)(uintptr_t)h, P(float*, out));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_mlp_forward(PyObject* self, PyObject* args) {
    unsigned long long h, pts, pl, po, out;
    int np, nc;
    if (!PyArg_ParseTuple(args, "KKiKKiK", &h, &pts, &np, &pl, &po, &nc, &out)) return NULL;
    NEED_CTX();
    int rc = gpu_mlp_forward(g_ctx, (gpu_mlp_t)(uintptr_t)h, P(const float*, pts), np, P(const float*, pl),
                             P(const int*, po), nc, P(float*, out));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_mlp_train_step(PyObject* self, PyObject* args) {
    unsigned long long h, pts, pl, po, tg, gr;
    int np, nc, apply;
    float lr;
    if (!PyArg_ParseTuple(args, "KKiKKiKfiK", &h, &pts, &np, &pl, &po, &nc, &tg, &lr, &apply, &gr)) return NULL;
    NEED_CTX();
    float loss = 0.0f;
    int rc = gpu_mlp_train_step(g_ctx, (gpu_mlp_t)(uintptr_t)h, P(const float*, pts), np, P(const float*, pl),
                                P(const int*, po), nc, P(const float*, tg), lr, apply, &loss,
                                gr ? P(float*, gr) : NULL);
    if (rc) return raise_error(rc);
    return PyFloat_FromDouble(loss);
}

static PyObject* m_mlp_sample_parts(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, out;
    int n;
    if (!PyArg_ParseTuple(args, "KKKKiK", &v, &vo, &t, &to, &n, &out)) return NULL;
    NEED_CTX();
    int rc = gpu_mlp_sample_parts(g_ctx, P(const float*, v), P(const int*, vo), P(const int*, t), P(const int*, to),
                                  n, P(float*, out));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_mlp_gemm_test(PyObject* self, PyObject* args) {
    int M, N, K;
    unsigned long long a, b, c;
    long long sam, sak, alen, sbk, sbn, blen;
    if (!PyArg_ParseTuple(args, "iiiKLLLKLLLK", &M, &N, &K, &a, &sam, &sak, &alen, &b, &sbk, &sbn, &blen, &c))
        return NULL;
    NEED_CTX();
    int rc = gpu_mlp_gemm_test(g_ctx, M, N, K, P(const unsigned short*, a), sam, sak, alen,
                               P(const unsigned short*, b), sbk, sbn, blen, P(float*, c));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_lookahead_set_proposal(PyObject* self, PyObject* args) {
    unsigned long long h;
    int k;
    if (!PyArg_ParseTuple(args, "Ki", &h, &k)) return NULL;
    NEED_CTX();
    int rc = gpu_lookahead_set_proposal(g_ctx, (gpu_mlp_t)(uintptr_t)h, k);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_lookahead_candidates(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, op, ok, oc;
    int n, width, nconc, cap;
    float thr, eps, cthr;
    if (!PyArg_ParseTuple(args, "iKKKKififfiKKK", &n, &v, &vo, &t, &to, &width, &thr, &nconc, &eps, &cthr, &cap, &op,
                          &ok, &oc))
        return NULL;
    NEED_CTX();
    int rc = gpu_lookahead_candidates(g_ctx, n, P(const float*, v), P(const int*, vo), P(const int*, t),
                                      P(const int*, to), width, thr, nconc, eps, cthr, cap, P(float*, op),
                                      P(int*, ok), P(int*, oc));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_lookahead_record(PyObject* self, PyObject* args) {
    int on;
    if (!PyArg_ParseTuple(args, "i", &on)) return NULL;
    NEED_CTX();
    int rc = gpu_lookahead_record(g_ctx, on);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_lookahead_record_size(PyObject* self, PyObject* args) {
    NEED_CTX();
    int np = 0, nc = 0;
    int rc = gpu_lookahead_record_size(g_ctx, &np, &nc);
    if (rc) return raise_error(rc);
    return Py_BuildValue("ii", np, nc);
}

static PyObject* m_lookahead_record_read(PyObject* self, PyObject* args) {
    unsigned long long a, b, c, d;
    if (!PyArg_ParseTuple(args, "KKKK", &a, &b, &c, &d)) return NULL;
    NEED_CTX();
    int rc = gpu_lookahead_record_read(g_ctx, P(float*, a), P(float*, b), P(int*, c), P(float*, d));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_copy_probe(PyObject* self, PyObject* args) {
    unsigned long long bytes;
    int reps;
    if (!PyArg_ParseTuple(args, "Ki", &bytes, &reps)) return NULL;
    NEED_CTX();
    float ms = 0.0f;
    int rc = gpu_copy_probe(g_ctx, (size_t)bytes, reps, &ms);
    if (rc) return raise_error(rc);
    return PyFloat_FromDouble(ms);
}

static PyObject* part_list(struct gpu_result* res) {
    PyObject* list = PyList_New(res->nparts);
    if (!list) return NULL;
    for (int i = 0; i < res->nparts; ++i) {
        struct gpu_part_result* p = &res->parts[i];
        PyObject* vb = PyBytes_FromStringAndSize((const char*)p->verts, (Py_ssize_t)p->nv * 12);
        PyObject* tb = PyBytes_FromStringAndSize((const char*)p->tris, (Py_ssize_t)p->nt * 12);
        PyObject* hvb = PyBytes_FromStringAndSize((const char*)p->hull_verts, (Py_ssize_t)p->hull_nv * 12);
        PyObject* htb = PyBytes_FromStringAndSize((const char*)p->hull_tris, (Py_ssize_t)p->hull_nt * 12);
        PyObject* tup = NULL;
        if (vb && tb && hvb && htb)
            tup = Py_BuildValue("(OOiiOOiifff)", vb, tb, p->nv, p->nt, hvb, htb, p->hull_nv, p->hull_nt,
                                (double)p->mesh_vol, (double)p->hull_vol, (double)p->hausdorff);
        Py_XDECREF(vb);
        Py_XDECREF(tb);
        Py_XDECREF(hvb);
        Py_XDECREF(htb);
        if (!tup) {
            Py_DECREF(list);
            return NULL;
        }
        PyList_SetItem(list, i, tup);
    }
    return list;
}

static PyObject* m_lookahead_batch(PyObject* self, PyObject* args) {
    unsigned long long v, vo, t, to, hv, hvo, ht, hto, outp, itp;
    int n, max_iters, width, width2, depth, quick, max_cut, dcomp, dc_iter, n_conc, conc_iters, merge;
    float thr, conc_eps, conc_thr;
    if (!PyArg_ParseTuple(args, "iKKKKKKKKiiifiiiiiiffiiKK", &n, &v, &vo, &t, &to, &hv, &hvo, &ht, &hto, &max_iters,
                          &width, &width2, &thr, &depth, &quick, &max_cut, &dcomp, &dc_iter, &n_conc, &conc_eps,
                          &conc_thr, &conc_iters, &merge, &outp, &itp))
        return NULL;
    NEED_CTX();
    struct gpu_result* res = (struct gpu_result*)calloc(n > 0 ? n : 1, sizeof(struct gpu_result));
    if (!res) return PyErr_NoMemory();
    int rc = gpu_lookahead_batch(g_ctx, n, P(const float*, v), P(const int*, vo), P(const int*, t), P(const int*, to),
                                 P(const float*, hv), P(const int*, hvo), P(const int*, ht), P(const int*, hto),
                                 max_iters, width, width2, thr, depth, quick, max_cut, dcomp, dc_iter, n_conc, conc_eps,
                                 conc_thr, conc_iters, merge, res, P(int*, itp), P(double*, outp));
    if (rc) {
        free(res);
        return raise_error(rc);
    }
    PyObject* all = PyList_New(n);
    for (int m = 0; m < n && all; ++m) {
        PyObject* l = part_list(&res[m]);
        if (!l) {
            Py_DECREF(all);
            all = NULL;
            break;
        }
        PyList_SetItem(all, m, l);
    }
    for (int m = 0; m < n; ++m) gpu_result_free(&res[m]);
    free(res);
    return all;
}

static PyObject* m_test_postprocess_dc(PyObject* self, PyObject* args) {
    unsigned long long vp, tp, ovp, otp, onvp, ontp;
    int nv, nt, mc, mv, mt;
    if (!PyArg_ParseTuple(args, "KiKiiiiKKKK", &vp, &nv, &tp, &nt, &mc, &mv, &mt, &ovp, &otp, &onvp, &ontp))
        return NULL;
    NEED_CTX();
    int ncomp = 0;
    int rc = gpu_test_postprocess_dc(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, mc, mv, mt,
                                     P(float*, ovp), P(int*, otp), P(int*, onvp), P(int*, ontp), &ncomp);
    if (rc) {
        PyErr_Format(PyExc_RuntimeError, "postprocess_dc failed: %s", gpu_last_error(g_ctx));
        return NULL;
    }
    return Py_BuildValue("i", ncomp);
}

static PyObject* m_lookahead_decompose(PyObject* self, PyObject* args, PyObject* kw) {
    static char* kwlist[] = {"verts_ptr", "nv", "tris_ptr", "nt", "hull_verts_ptr", "hull_nv", "hull_tris_ptr",
                             "hull_nt", "max_iters", "width", "width2", "threshold", "depth", "quick_depth",
                             "max_n_cutting", "verbose", "debug", "decompose_components",
                             "no_decompose_components_per_iter", "n_concave_edges", "concave_eps",
                             "concave_threshold", "concave_iters", "merge_hulls", NULL};
    unsigned long long vp, tp, hvp, htp;
    int nv, nt, hnv, hnt, max_iters, width, width2;
    float threshold;
    int depth = 2, quick_depth = 1, max_n_cutting = 16, verbose = 0, debug = 0, dcomp = 0, no_dc_per_iter = 0;
    int n_concave = 32, concave_iters = 10, merge = 0;
    float concave_eps = 0.005f, concave_thr = 3.49f;
    if (!PyArg_ParseTupleAndKeywords(args, kw, "KiKiKiKiiiifiii|iiiiiffii", kwlist, &vp, &nv, &tp, &nt, &hvp, &hnv,
                                     &htp, &hnt, &max_iters, &width, &width2, &threshold, &depth, &quick_depth,
                                     &max_n_cutting, &verbose, &debug, &dcomp, &no_dc_per_iter, &n_concave,
                                     &concave_eps, &concave_thr, &concave_iters, &merge))
        return NULL;
    NEED_CTX();
    struct gpu_result res;
    memset(&res, 0, sizeof(res));
    int rc = lookahead_decompose(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, P(const float*, hvp), hnv,
                                 P(const int*, htp), hnt, max_iters, width, width2, threshold, depth, quick_depth,
                                 max_n_cutting, verbose, debug, dcomp, !no_dc_per_iter, n_concave, concave_eps,
                                 concave_thr, concave_iters, merge, &res);
    if (rc) {
        gpu_result_free(&res);
        return raise_error(rc);
    }
    PyObject* list = PyList_New(res.nparts);
    if (!list) {
        gpu_result_free(&res);
        return NULL;
    }
    for (int i = 0; i < res.nparts; ++i) {
        struct gpu_part_result* p = &res.parts[i];
        PyObject* vb = PyBytes_FromStringAndSize((const char*)p->verts, (Py_ssize_t)p->nv * 12);
        PyObject* tb = PyBytes_FromStringAndSize((const char*)p->tris, (Py_ssize_t)p->nt * 12);
        PyObject* hvb = PyBytes_FromStringAndSize((const char*)p->hull_verts, (Py_ssize_t)p->hull_nv * 12);
        PyObject* htb = PyBytes_FromStringAndSize((const char*)p->hull_tris, (Py_ssize_t)p->hull_nt * 12);
        PyObject* tup = NULL;
        if (vb && tb && hvb && htb)
            tup = Py_BuildValue("(OOiiOOiiff)", vb, tb, p->nv, p->nt, hvb, htb, p->hull_nv, p->hull_nt,
                                (double)p->mesh_vol, (double)p->hull_vol);
        Py_XDECREF(vb);
        Py_XDECREF(tb);
        Py_XDECREF(hvb);
        Py_XDECREF(htb);
        if (!tup) {
            Py_DECREF(list);
            gpu_result_free(&res);
            return NULL;
        }
        PyList_SetItem(list, i, tup);
    }
    gpu_result_free(&res);
    return list;
}

static PyObject* m_kdop_hull(PyObject* self, PyObject* args) {
    unsigned long long pp, op, vp, tp, nvp, ntp, volp, ep;
    int total, nh, mhv, mht;
    if (!PyArg_ParseTuple(args, "KiKiiiKKKKKK", &pp, &total, &op, &nh, &mhv, &mht, &vp, &tp, &nvp, &ntp, &volp,
                          &ep))
        return NULL;
    NEED_CTX();
    int rc = gpu_kdop_hull(g_ctx, P(const float*, pp), total, P(const int*, op), nh, mhv, mht, P(float*, vp),
                           P(int*, tp), P(int*, nvp), P(int*, ntp), P(float*, volp), P(int*, ep));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_test_hausdorff(PyObject* self, PyObject* args) {
    unsigned long long hv, ht, mv, mt, op;
    int hnv, hnt, mnv, mnt;
    if (!PyArg_ParseTuple(args, "KiKiKiKiK", &hv, &hnv, &ht, &hnt, &mv, &mnv, &mt, &mnt, &op)) return NULL;
    NEED_CTX();
    int rc = gpu_test_hausdorff(g_ctx, P(const float*, hv), hnv, P(const int*, ht), hnt, P(const float*, mv), mnv,
                                P(const int*, mt), mnt, P(float*, op));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}

static PyObject* m_mesh_audit(PyObject* self, PyObject* args) {
    unsigned long long vp, tp;
    int nv, nt;
    if (!PyArg_ParseTuple(args, "KiKi", &vp, &nv, &tp, &nt)) return NULL;
    NEED_CTX();
    unsigned int flags = 0;
    int rc = gpu_mesh_audit(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, &flags);
    if (rc) return raise_error(rc);
    return PyLong_FromUnsignedLong(flags);
}

static PyObject* m_preprocess(PyObject* self, PyObject* args) {
    unsigned long long vp, tp;
    int nv, nt, res;
    if (!PyArg_ParseTuple(args, "KiKii", &vp, &nv, &tp, &nt, &res)) return NULL;
    NEED_CTX();
    float* ov = NULL;
    int* ot = NULL;
    int onv = 0, ont = 0;
    int rc = gpu_preprocess(g_ctx, P(const float*, vp), nv, P(const int*, tp), nt, res, &ov, &onv, &ot, &ont);
    if (rc) {
        free(ov);
        free(ot);
        return raise_error(rc);
    }
    PyObject* vb = PyBytes_FromStringAndSize((const char*)ov, (Py_ssize_t)onv * 12);
    PyObject* tb = PyBytes_FromStringAndSize((const char*)ot, (Py_ssize_t)ont * 12);
    free(ov);
    free(ot);
    if (!vb || !tb) {
        Py_XDECREF(vb);
        Py_XDECREF(tb);
        return NULL;
    }
    PyObject* tup = Py_BuildValue("(OOii)", vb, tb, onv, ont);
    Py_DECREF(vb);
    Py_DECREF(tb);
    return tup;
}

/* ---- device-resident plumbing (no reference counterpart) ------------------ */
static PyObject* m_dev_alloc(PyObject* self, PyObject* args) {
    unsigned long long n;
    if (!PyArg_ParseTuple(args, "K", &n)) return NULL;
    NEED_CTX();
    void* p = NULL;
    int rc = gpu_dev_alloc(g_ctx, (size_t)n, &p);
    if (rc) return raise_error(rc);
    return PyLong_FromUnsignedLongLong((unsigned long long)(uintptr_t)p);
}
static PyObject* m_dev_free(PyObject* self, PyObject* args) {
    unsigned long long p;
    if (!PyArg_ParseTuple(args, "K", &p)) return NULL;
    NEED_CTX();
    int rc = gpu_dev_free(g_ctx, P(void*, p));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_h2d(PyObject* self, PyObject* args) {
    unsigned long long d, s, n;
    if (!PyArg_ParseTuple(args, "KKK", &d, &s, &n)) return NULL;
    NEED_CTX();
    int rc = gpu_memcpy_h2d(g_ctx, P(void*, d), P(const void*, s), (size_t)n);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_d2h(PyObject* self, PyObject* args) {
    unsigned long long d, s, n;
    if (!PyArg_ParseTuple(args, "KKK", &d, &s, &n)) return NULL;
    NEED_CTX();
    int rc = gpu_memcpy_d2h(g_ctx, P(void*, d), P(const void*, s), (size_t)n);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_sync(PyObject* self, PyObject* args) {
    NEED_CTX();
    int rc = gpu_sync(g_ctx);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_timer_start(PyObject* self, PyObject* args) {
    NEED_CTX();
    int rc = gpu_timer_start(g_ctx);
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_timer_stop(PyObject* self, PyObject* args) {
    NEED_CTX();
    float ms = 0.f;
    int rc = gpu_timer_stop(g_ctx, &ms);
    if (rc) return raise_error(rc);
    return PyFloat_FromDouble((double)ms);
}
static PyObject* m_device_info(PyObject* self, PyObject* args) {
    NEED_CTX();
    char name[256];
    int ncu = 0;
    size_t hbm = 0;
    int rc = gpu_device_info(g_ctx, name, (int)sizeof(name), &ncu, &hbm);
    if (rc) return raise_error(rc);
    return Py_BuildValue("(siK)", name, ncu, (unsigned long long)hbm);
}
/* hull_dandc_dev(d_pts, d_offsets, n_hulls, max_pts, max_hv, max_ht, d_verts, d_tris, d_nv, d_nt, d_err, d_vol) */
static PyObject* m_hull_dandc_dev(PyObject* self, PyObject* args) {
    unsigned long long pp, op, vp, tp, nvp, ntp, ep, volp;
    int nh, mp, mhv, mht;
    if (!PyArg_ParseTuple(args, "KKiiiiKKKKKK", &pp, &op, &nh, &mp, &mhv, &mht, &vp, &tp, &nvp, &ntp, &ep, &volp))
        return NULL;
    NEED_CTX();
    int rc = gpu_hull_dandc_dev(g_ctx, P(const float*, pp), P(const int*, op), nh, mp, mhv, mht, P(float*, vp),
                                P(int*, tp), P(int*, nvp), P(int*, ntp), P(int*, ep), P(double*, volp));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
/* hausdorff_batch(av, avoff, at, atoff, bv, bvoff, bt, btoff, n_pairs, mode, n_samples,
 *                 out_ab, out_ba, samples_a, samples_b)  (0 for a NULL pointer) */
static PyObject* m_hausdorff_batch(PyObject* self, PyObject* args) {
    unsigned long long av, avo, at, ato, bv, bvo, bt, bto, oab, oba, sa, sb;
    int n, mode, ns;
    if (!PyArg_ParseTuple(args, "KKKKKKKKiiiKKKK", &av, &avo, &at, &ato, &bv, &bvo, &bt, &bto, &n, &mode, &ns, &oab,
                          &oba, &sa, &sb))
        return NULL;
    NEED_CTX();
    int rc = gpu_hausdorff_batch(g_ctx, P(const float*, av), P(const int*, avo), P(const int*, at),
                                 P(const int*, ato), P(const float*, bv), P(const int*, bvo), P(const int*, bt),
                                 P(const int*, bto), n, mode, ns, P(float*, oab), P(float*, oba), P(float*, sa),
                                 P(float*, sb));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_hausdorff_plan_create(PyObject* self, PyObject* args) {
    unsigned long long av, at, bv, bt, avo, ato, bvo, bto, oab, oba, sa, sb;
    int n, mode, ns;
    if (!PyArg_ParseTuple(args, "KKKKKKKKiiiKKKK", &av, &at, &bv, &bt, &avo, &ato, &bvo, &bto, &n, &mode, &ns, &oab,
                          &oba, &sa, &sb))
        return NULL;
    NEED_CTX();
    hd_plan_t plan = NULL;
    int rc = gpu_hausdorff_plan_create(g_ctx, P(const float*, av), P(const int*, at), P(const float*, bv),
                                       P(const int*, bt), P(const int*, avo), P(const int*, ato),
                                       P(const int*, bvo), P(const int*, bto), n, mode, ns, P(float*, oab),
                                       P(float*, oba), P(float*, sa), P(float*, sb), &plan);
    if (rc) return raise_error(rc);
    return PyLong_FromUnsignedLongLong((unsigned long long)(uintptr_t)plan);
}
static PyObject* m_hausdorff_plan_run(PyObject* self, PyObject* args) {
    unsigned long long p;
    if (!PyArg_ParseTuple(args, "K", &p)) return NULL;
    NEED_CTX();
    int rc = gpu_hausdorff_plan_run(g_ctx, P(hd_plan_t, p));
    if (rc) return raise_error(rc);
    Py_RETURN_NONE;
}
static PyObject* m_hausdorff_plan_destroy(PyObject* self, PyObject* args) {
    unsigned long long p;
    if (!PyArg_ParseTuple(args, "K", &p)) return NULL;
    gpu_hausdorff_plan_destroy(g_ctx, P(hd_plan_t, p));
    Py_RETURN_NONE;
}

static PyObject* m_inject_error(PyObject* self, PyObject* args) {
    int bits;
    if (!PyArg_ParseTuple(args, "i", &bits)) return NULL;
    NEED_CTX();
    gpu_debug_inject_error(g_ctx, bits);
    Py_RETURN_NONE;
}

static PyMethodDef methods[] = {
    {"set_fatbin", m_set_fatbin, METH_O, "Accepts and ignores a kernel blob (kernels are linked in)."},
    {"init", (PyCFunction)(void (*)(void))m_init, METH_VARARGS | METH_KEYWORDS, "init(device=-1, pool_bytes=0)"},
    {"destroy", m_destroy, METH_NOARGS, "Destroy the GPU context."},
    {"heap_compact", m_heap_compact, METH_NOARGS, "Synchronise the device heaps."},
    {"pool_usage", m_pool_usage, METH_NOARGS, "Bytes bumped from the device pool (monotonic)."},
    {"heap_stats", m_heap_stats, METH_NOARGS, "(heap_out, heap_allocs, heap_frees, scr_out, scr_allocs, scr_frees)"},
    {"test_warp_sort", m_test_warp_sort, METH_VARARGS, "Sort (x,y,z,index) records by (y,x,z,index)."},
    {"hull_dandc", m_hull_dandc, METH_VARARGS, "Exact hull mesh per point cloud."},
    {"hull_dandc_vol", m_hull_dandc_vol, METH_VARARGS, "hull_dandc plus FP64 volumes."},
    {"test_mesh_volume", m_test_mesh_volume, METH_VARARGS, "Volume of one mesh."},
    {"batch_mesh_volume", m_batch_mesh_volume, METH_VARARGS, "Volumes of a batch of meshes."},
    {"test_plane_cut", m_test_plane_cut, METH_VARARGS, "Plane cut with cap triangulation."},
    {"plane_cut_batch", m_plane_cut_batch, METH_VARARGS, "Batched plane cuts: volumes, sizes, status."},
    {"mlp_n_params", m_mlp_n_params, METH_NOARGS, "Parameter count of the proposal network (padded layout)."},
    {"mlp_create", m_mlp_create, METH_VARARGS, "Create the proposal network (params pointer or 0, seed)."},
    {"mlp_destroy", m_mlp_destroy, METH_VARARGS, "Destroy a proposal network."},
    {"mlp_get_params", m_mlp_get_params, METH_VARARGS, "Copy the FP32 parameters out."},
    {"mlp_forward", m_mlp_forward, METH_VARARGS, "Predicted scores of candidate planes."},
    {"mlp_train_step", m_mlp_train_step, METH_VARARGS, "Forward, backward and Adam step; returns the loss."},
    {"mlp_sample_parts", m_mlp_sample_parts, METH_VARARGS, "Surface samples of parts in their box frames."},
    {"mlp_gemm_test", m_mlp_gemm_test, METH_VARARGS, "The network's BF16 MFMA GEMM on host operands."},
    {"lookahead_record", m_lookahead_record, METH_VARARGS, "Start (1) or discard (0) the level-0 training record."},
    {"lookahead_record_size", m_lookahead_record_size, METH_NOARGS, "(parts, candidates) recorded."},
    {"lookahead_record_read", m_lookahead_record_read, METH_VARARGS, "Copy the record out."},
    {"lookahead_candidates", m_lookahead_candidates, METH_VARARGS, "Level-0 candidate planes of whole meshes."},
    {"lookahead_set_proposal", m_lookahead_set_proposal, METH_VARARGS, "Learned proposal for the lookahead."},
    {"copy_probe", m_copy_probe, METH_VARARGS, "Known-size device copy (bytes, reps) -> ms per copy."},
    {"set_tuning", m_set_tuning, METH_VARARGS, "Set a hull launch-shape knob (key, value)."},
    {"get_tuning", m_get_tuning, METH_VARARGS, "Read a hull launch-shape knob."},
    {"score_cuts", m_score_cuts, METH_VARARGS, "Batched exact cut scores: cut, hulls, max(0.3 rv, H) per half."},
    {"score_plan_create", m_score_plan_create, METH_VARARGS, "Plan a batch of cut scores as one captured graph."},
    {"score_plan_run", m_score_plan_run, METH_VARARGS, "Replay a score plan; scores and detail out."},
    {"score_plan_destroy", m_score_plan_destroy, METH_VARARGS, "Free a score plan."},
    {"kdop_hull", m_kdop_hull, METH_VARARGS, "k-DOP prefiltered exact hull."},
    {"test_hausdorff", m_test_hausdorff, METH_VARARGS, "Bidirectional sampled Hausdorff distance."},
    {"test_postprocess_dc", m_test_postprocess_dc, METH_VARARGS, "Split a mesh into connected components."},
    {"mesh_audit", m_mesh_audit, METH_VARARGS, "Topology audit -> MAV_* bitmask."},
    {"preprocess", m_preprocess, METH_VARARGS, "Not part of this library (raises)."},
    {"lookahead_decompose", (PyCFunction)(void (*)(void))m_lookahead_decompose, METH_VARARGS | METH_KEYWORDS,
     "Lookahead tree-search convex decomposition."},
    {"lookahead_batch", m_lookahead_batch, METH_VARARGS, "Batched lookahead decomposition of packed meshes."},
    {"dev_alloc", m_dev_alloc, METH_VARARGS, "Device allocation -> address."},
    {"dev_free", m_dev_free, METH_VARARGS, "Free a device allocation."},
    {"memcpy_h2d", m_h2d, METH_VARARGS, "Host to device copy (dst, src, bytes)."},
    {"memcpy_d2h", m_d2h, METH_VARARGS, "Device to host copy (dst, src, bytes)."},
    {"sync", m_sync, METH_NOARGS, "Synchronise the context stream."},
    {"timer_start", m_timer_start, METH_NOARGS, "Record the start event on the context stream."},
    {"timer_stop", m_timer_stop, METH_NOARGS, "Record the stop event, wait, return elapsed ms."},
    {"device_info", m_device_info, METH_NOARGS, "(name, n_cu, hbm_bytes)"},
    {"hull_dandc_dev", m_hull_dandc_dev, METH_VARARGS, "Device-pointer hull batch (stream-ordered)."},
    {"inject_error", m_inject_error, METH_VARARGS, "OR KERR bits into the next error readback."},
    {"hausdorff_batch", m_hausdorff_batch, METH_VARARGS, "Directed sampled Hausdorff distances for mesh pairs."},
    {"hausdorff_plan_create", m_hausdorff_plan_create, METH_VARARGS, "Plan a device-resident Hausdorff batch."},
    {"hausdorff_plan_run", m_hausdorff_plan_run, METH_VARARGS, "Launch a Hausdorff plan (stream-ordered)."},
    {"hausdorff_plan_destroy", m_hausdorff_plan_destroy, METH_VARARGS, "Free a Hausdorff plan."},
    {NULL, NULL, 0, NULL}};

static int exec_module(PyObject* m) { return PyModule_AddFunctions(m, methods); }

static PyModuleDef_Slot slots[] = {{Py_mod_exec, (void*)exec_module}, {0, NULL}};

static PyModuleDef moddef = {PyModuleDef_HEAD_INIT, "cuacd_amd._gpu", "MI355X convex decomposition kernels", 0,
                             NULL, slots, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__gpu(void) { return PyModuleDef_Init(&moddef); }
