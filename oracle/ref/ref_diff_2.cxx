/*
 * ref_diff_2.cxx -- wraps the reference's own src/diff_2.cxx kernels.
 * TEST INFRASTRUCTURE ONLY: builds oracle/_ref, which writes golden vectors.
 */
#include "diff_2.cxx"
#include "ref_io.h"

namespace
{
    void op_c(const Ref_run& r)
    {
        Ref_dims d(r);
        auto at = r.load("at", d.ncells); auto a = r.load("a", d.ncells);
        auto dzi = r.load("dzi", d.kcells); auto dzhi = r.load("dzhi", d.kcells);
        diff_c<double>(at.data(), a.data(), r.p("visc"), d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend,
                       d.icells, d.ijcells, d.dx, d.dy, dzi.data(), dzhi.data());
        r.save("at", at);
    }
    void op_w(const Ref_run& r)
    {
        Ref_dims d(r);
        auto at = r.load("at", d.ncells); auto a = r.load("a", d.ncells);
        auto dzi = r.load("dzi", d.kcells); auto dzhi = r.load("dzhi", d.kcells);
        diff_w<double>(at.data(), a.data(), r.p("visc"), d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend,
                       d.icells, d.ijcells, d.dx, d.dy, dzi.data(), dzhi.data());
        r.save("at", at);
    }
}

void ref_register_diff_2()
{
    ref_register("diff_2_c", op_c);
    ref_register("diff_2_w", op_w);
}
