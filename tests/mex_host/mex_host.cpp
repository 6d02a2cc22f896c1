// Test host of the MATLAB gateway (see mex.h): mxArray values, and C entry
// points through which Python builds the inputs, calls mexFunction and
// reads the output struct.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

struct mxArray_tag {
    bool is_struct = false, is_complex = false;
    size_t m = 0, n = 0;
    std::vector<double> re, im;
    std::vector<std::string> fields;
    std::vector<std::map<std::string, mxArray *>> elems;   // struct elements
};

namespace {
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
} // namespace

extern "C" {

mxArray *mxGetField(const mxArray *s, mwIndex idx, const char *name)
{
    if (!s || !s->is_struct || idx >= s->elems.size()) return nullptr;
    auto it = s->elems[idx].find(name);
    return it == s->elems[idx].end() ? nullptr : it->second;
}

void mxSetField(mxArray *s, mwIndex idx, const char *name, mxArray *v)
{
    if (!s->is_struct || idx >= s->elems.size()) throw MexError("mxSetField out of range");
    s->elems[idx][name] = v;
}

bool mxIsEmpty(const mxArray *a) { return a->m == 0 || a->n == 0; }
bool mxIsStruct(const mxArray *a) { return a && a->is_struct; }
bool mxIsComplex(const mxArray *a) { return a->is_complex; }
double mxGetScalar(const mxArray *a) { return a->re.empty() ? 0.0 : a->re[0]; }
double *mxGetPr(const mxArray *a) { return const_cast<double *>(a->re.data()); }
double *mxGetPi(const mxArray *a) { return a->is_complex ? const_cast<double *>(a->im.data()) : nullptr; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    mxArray *a = new mxArray_tag();
    a->m = m;
    a->n = n;
    a->re.assign(m * n, 0.0);
    a->is_complex = c == mxCOMPLEX;
    if (a->is_complex) a->im.assign(m * n, 0.0);
    return a;
}

mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    a->re[0] = v;
    return a;
}

mxArray *mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char **names)
{
    mxArray *a = new mxArray_tag();
    a->is_struct = true;
    a->m = m;
    a->n = n;
    for (int i = 0; i < nfields; ++i) a->fields.push_back(names[i]);
    a->elems.resize(m * n);
    return a;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MexError(std::string(id) + ": " + buf);
}

// ---- entry points for the test ----

mxArray *mh_double(size_t m, size_t n, const double *re, const double *im)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, im ? mxCOMPLEX : mxREAL);
    for (size_t i = 0; i < m * n; ++i) {
        a->re[i] = re[i];
        if (im) a->im[i] = im[i];
    }
    return a;
}

mxArray *mh_struct(size_t n) { return mxCreateStructMatrix(1, n, 0, nullptr); }
void mh_set(mxArray *s, size_t idx, const char *name, mxArray *v) { mxSetField(s, idx, name, v); }
mxArray *mh_get(const mxArray *s, size_t idx, const char *name) { return mxGetField(s, idx, name); }

// shape: m, n, complex flag, struct flag
void mh_shape(const mxArray *a, size_t *m, size_t *n, int *cplx, int *strct)
{
    *m = a->m;
    *n = a->n;
    *cplx = a->is_complex;
    *strct = a->is_struct;
}

void mh_data(const mxArray *a, double *re, double *im)
{
    for (size_t i = 0; i < a->re.size(); ++i) {
        re[i] = a->re[i];
        if (im && a->is_complex) im[i] = a->im[i];
    }
}

void mh_free(mxArray *a)
{
    if (!a) return;
    for (auto &e : a->elems)
        for (auto &kv : e) mh_free(kv.second);
    delete a;
}

// calls the gateway; returns 0, or 1 with the error text in err
int mh_call(int nrhs, const mxArray **prhs, mxArray **out, char *err, size_t errlen)
{
    mxArray *plhs[1] = {nullptr};
    try {
        mexFunction(1, plhs, nrhs, prhs);
    } catch (const std::exception &e) {
        std::snprintf(err, errlen, "%s", e.what());
        return 1;
    }
    *out = plhs[0];
    return 0;
}

} // extern "C"
