// ops_seq_v2.h -- the host-side subset of the OPS API that MPLB Apps call
// directly (ops_init, ops_par_loop over ACC<T> accessors, ops_arg_dat/gbl/
// idx, ops_printf, ops_timers, ops_timing_output, ops_diagnostic_output,
// ops_exit).  Loops run on the host over host mirrors of the HiLeMMS
// fields; the mirrors are synchronised with the device by the HiLeMMS
// layer (a field written here is uploaded before the next device use).
// There is no generated code and no device loop driver behind this header:
// the lattice update itself runs in libmplbhip.
#ifndef OPS_SEQ_V2_H
#define OPS_SEQ_V2_H

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <ostream>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#define OPS_READ 0
#define OPS_WRITE 1
#define OPS_RW 2
#define OPS_INC 3
#define OPS_MIN 4
#define OPS_MAX 5

struct ops_block_core {
    int dims;
    std::string name;
};
typedef ops_block_core* ops_block;

// A field of one block: AoS storage with `dim` components per node and a
// halo of depth d on every space axis, i fastest (OPS' default layout).
struct ops_dat_core {
    std::string name, type;
    int block_index{0};
    int dim{1};
    int size[3]{1, 1, 1};
    int d_m[3]{0, 0, 0};
    int d_p[3]{0, 0, 0};
    size_t elem_size{8};
    std::vector<char> data;
    bool dirty{false};  // written on the host since the last device sync
    long long nodes() const {
        return (long long)(size[0] - d_m[0] + d_p[0]) *
               (size[1] - d_m[1] + d_p[1]) * (size[2] - d_m[2] + d_p[2]);
    }
    long long stride(int a) const {
        long long s = dim;
        for (int i = 0; i < a; i++) s *= (size[i] - d_m[i] + d_p[i]);
        return s;
    }
    // element offset of node (i,j,k), component 0
    long long offset(int i, int j, int k) const {
        return (i - d_m[0]) * stride(0) + (j - d_m[1]) * stride(1) +
               (k - d_m[2]) * stride(2);
    }
};
typedef ops_dat_core* ops_dat;

struct ops_stencil_core {
    int dims{0};
    int points{0};
    std::string name;
};
typedef ops_stencil_core* ops_stencil;

struct ops_reduction_core {
    std::string name;
    double value{0};
};
typedef ops_reduction_core* ops_reduction;

enum ops_arg_kind { OPS_ARG_DAT, OPS_ARG_GBL, OPS_ARG_IDX, OPS_ARG_REDUCE };
struct ops_arg {
    ops_arg_kind kind;
    ops_dat dat{nullptr};
    void* ptr{nullptr};
    int dim{1};
    int acc{OPS_READ};
};

ops_stencil ops_decl_stencil(int dims, int points, int* sten, const char* name);
ops_block ops_decl_block(int dims, const char* name);
ops_dat ops_decl_dat_host(ops_block block, int block_index, int dim, const int* size,
                          int halo, size_t elem_size, const char* type,
                          const char* name);
ops_reduction ops_decl_reduction_handle(int size, const char* type, const char* name);
void ops_reduction_result(ops_reduction h, double* out);

inline ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil, const char*, int acc) {
    ops_arg a;
    a.kind = OPS_ARG_DAT;
    a.dat = dat;
    a.dim = dim;
    a.acc = acc;
    if (acc != OPS_READ) dat->dirty = true;
    return a;
}
template <typename T>
inline ops_arg ops_arg_gbl(T* ptr, int dim, const char*, int acc) {
    ops_arg a;
    a.kind = OPS_ARG_GBL;
    a.ptr = (void*)ptr;
    a.dim = dim;
    a.acc = acc;
    return a;
}
inline ops_arg ops_arg_idx() {
    ops_arg a;
    a.kind = OPS_ARG_IDX;
    return a;
}
inline ops_arg ops_arg_reduce(ops_reduction h, int dim, const char*, int acc) {
    ops_arg a;
    a.kind = OPS_ARG_REDUCE;
    a.ptr = (void*)h;
    a.dim = dim;
    a.acc = acc;
    return a;
}

// Accessor of one node of a field: acc(i,j[,k]) for scalar fields,
// acc(d,i,j[,k]) for vector fields (OPS v2 ACC<> semantics).
template <typename T>
class ACC {
   public:
    ACC() = default;
    ACC(T* p, long long s0, long long s1, long long s2, int dim)
        : p_(p), s_{s0, s1, s2}, dim_(dim) {}
#ifdef OPS_3D
    T& operator()(int i, int j, int k) const { return p_[i * s_[0] + j * s_[1] + k * s_[2]]; }
    T& operator()(int d, int i, int j, int k) const {
        return p_[d + i * s_[0] + j * s_[1] + k * s_[2]];
    }
#else
    T& operator()(int i, int j) const { return p_[i * s_[0] + j * s_[1]]; }
    T& operator()(int d, int i, int j) const { return p_[d + i * s_[0] + j * s_[1]]; }
#endif
   private:
    T* p_{nullptr};
    long long s_[3]{0, 0, 0};
    int dim_{1};
};

namespace ops_host_detail {
template <typename P>
struct Param {
    using Bare = typename std::remove_cv<typename std::remove_reference<P>::type>::type;
};
template <typename T>
struct IsAcc : std::false_type {};
template <typename T>
struct IsAcc<ACC<T>> : std::true_type {
    using Elem = T;
};

template <typename P>
auto make_param(const ops_arg& a, const int* idx, int* idxbuf) {
    using Bare = typename Param<P>::Bare;
    if constexpr (IsAcc<Bare>::value) {
        using E = typename IsAcc<Bare>::Elem;
        ops_dat d = a.dat;
        E* base = reinterpret_cast<E*>(d->data.data()) + d->offset(idx[0], idx[1], idx[2]);
        return Bare(base, d->stride(0), d->stride(1), d->stride(2), d->dim);
    } else {
        // pointers: globals, reductions and the index argument
        if (a.kind == OPS_ARG_IDX) {
            for (int i = 0; i < 3; i++) idxbuf[i] = idx[i];
            return Bare(idxbuf);
        }
        if (a.kind == OPS_ARG_REDUCE)
            return Bare(&static_cast<ops_reduction>(a.ptr)->value);
        return Bare(a.ptr);
    }
}

template <typename... P, typename... A, size_t... I>
void call_node(void (*kernel)(P...), const std::tuple<A...>& args, const int* idx,
               std::index_sequence<I...>) {
    int idxbuf[3];
    auto params = std::make_tuple(make_param<P>(std::get<I>(args), idx, idxbuf)...);
    kernel(std::get<I>(params)...);
}
}  // namespace ops_host_detail

// ops_par_loop over range = {x0,x1,y0,y1[,z0,z1]}, i fastest.
template <typename... P, typename... A>
void ops_par_loop(void (*kernel)(P...), const char* /*name*/, ops_block /*block*/,
                  int dim, int* range, A... args) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel/argument count mismatch");
    const std::tuple<A...> targs(args...);
    const int z0 = dim == 3 ? range[4] : 0, z1 = dim == 3 ? range[5] : 1;
    int idx[3];
    for (int k = z0; k < z1; k++)
        for (int j = range[2]; j < range[3]; j++)
            for (int i = range[0]; i < range[1]; i++) {
                idx[0] = i;
                idx[1] = j;
                idx[2] = k;
                ops_host_detail::call_node(kernel, targs, idx,
                                           std::index_sequence_for<P...>{});
            }
}

void ops_init(const int argc, const char** argv, const int diags_level);
void ops_exit();
void ops_printf(const char* format, ...);
void ops_timers(double* cpu, double* et);
void ops_timing_output(std::ostream& stream);
void ops_diagnostic_output();
void ops_partition(const char* routine);
void ops_decl_const(const char* name, int dim, const char* type, void* data);

#endif
