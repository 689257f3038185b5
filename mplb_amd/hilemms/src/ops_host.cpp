// ops_host.cpp -- host side of the OPS subset in ops_seq_v2.h: blocks,
// host-mirror dats, stencil names, timers and printing.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "ops_seq_v2.h"

namespace {
std::vector<std::unique_ptr<ops_block_core>> g_blocks;
std::vector<std::unique_ptr<ops_dat_core>> g_dats;
std::vector<std::unique_ptr<ops_stencil_core>> g_stencils;
std::vector<std::unique_ptr<ops_reduction_core>> g_reductions;
std::chrono::steady_clock::time_point g_t0 = std::chrono::steady_clock::now();
}  // namespace

ops_stencil ops_decl_stencil(int dims, int points, int*, const char* name) {
    g_stencils.emplace_back(new ops_stencil_core());
    g_stencils.back()->dims = dims;
    g_stencils.back()->points = points;
    g_stencils.back()->name = name ? name : "";
    return g_stencils.back().get();
}

ops_block ops_decl_block(int dims, const char* name) {
    g_blocks.emplace_back(new ops_block_core());
    g_blocks.back()->dims = dims;
    g_blocks.back()->name = name ? name : "";
    return g_blocks.back().get();
}

ops_dat ops_decl_dat_host(ops_block block, int block_index, int dim, const int* size,
                          int halo, size_t elem_size, const char* type, const char* name) {
    g_dats.emplace_back(new ops_dat_core());
    ops_dat d = g_dats.back().get();
    d->name = name ? name : "";
    d->type = type ? type : "double";
    d->block_index = block_index;
    d->dim = dim;
    const int dims = block ? block->dims : 3;
    for (int a = 0; a < 3; a++) {
        d->size[a] = size[a];
        d->d_m[a] = a < dims ? -halo : 0;
        d->d_p[a] = a < dims ? halo : 0;
    }
    d->elem_size = elem_size;
    // zero-filled, as OPS creates a dat without data
    d->data.assign((size_t)d->nodes() * dim * elem_size, 0);
    return d;
}

ops_reduction ops_decl_reduction_handle(int, const char*, const char* name) {
    g_reductions.emplace_back(new ops_reduction_core());
    g_reductions.back()->name = name ? name : "";
    return g_reductions.back().get();
}

void ops_reduction_result(ops_reduction h, double* out) {
    *out = h->value;
    h->value = 0;
}

void ops_init(const int, const char**, const int) { g_t0 = std::chrono::steady_clock::now(); }
void ops_exit() {}
void ops_partition(const char*) {}
void ops_decl_const(const char*, int, const char*, void*) {}

void ops_printf(const char* format, ...) {
    va_list ap;
    va_start(ap, format);
    std::vprintf(format, ap);
    va_end(ap);
    std::fflush(stdout);
}

void ops_timers(double* cpu, double* et) {
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - g_t0).count();
    if (cpu) *cpu = s;
    if (et) *et = s;
}

void ops_timing_output(std::ostream& stream) {
    stream << "lattice update timings: see rocprofv3 --kernel-trace --stats "
              "(kernel k_lattice)\n";
}

void ops_diagnostic_output() {
    ops_printf("host dats: %zu, blocks: %zu\n", g_dats.size(), g_blocks.size());
}
