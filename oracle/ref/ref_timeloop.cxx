/*
 * ref_timeloop.cxx -- wraps the reference's own src/timeloop.cxx rk3 kernel.
 * TEST INFRASTRUCTURE ONLY: builds oracle/_ref, which writes golden vectors.
 */
#include "timeloop.cxx"
#include "ref_io.h"

namespace
{
    void op_rk3(const Ref_run& r)
    {
        Ref_dims d(r);
        auto a = r.load("a", d.ncells); auto at = r.load("at", d.ncells);
        rk3<double>(a.data(), at.data(), r.pi("substep"), r.p("dt"), d.istart, d.iend, d.jstart, d.jend,
                    d.kstart, d.kend, d.icells, d.ijcells, d.ncells);
        r.save("a", a);
        r.save("at", at);
    }
}

void ref_register_timeloop()
{
    ref_register("rk3", op_rk3);
}
