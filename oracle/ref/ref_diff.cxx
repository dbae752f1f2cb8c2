/*
 * ref_diff.cxx -- wraps the reference's header-only smag2 kernels
 * (include/diff_kernels.h).
 * TEST INFRASTRUCTURE ONLY: builds oracle/_ref, which writes golden vectors.
 */
#include <vector>
#include "diff_kernels.h"
#include "ref_io.h"

namespace
{
    namespace dk = Diff_kernels;

    template<Surface_model sm>
    void strain2(const Ref_run& r)
    {
        Ref_dims d(r); const size_t n = d.ncells, nk = d.kcells, n2 = d.ijcells;
        auto s2 = r.load("strain2", n); auto u = r.load("u", n); auto v = r.load("v", n); auto w = r.load("w", n);
        auto ug = r.load("ugradbot", n2); auto vg = r.load("vgradbot", n2);
        auto z = r.load("z", nk); auto dzi = r.load("dzi", nk); auto dzhi = r.load("dzhi", nk);
        dk::calc_strain2<double, sm>(s2.data(), u.data(), v.data(), w.data(), ug.data(), vg.data(),
                z.data(), dzi.data(), dzhi.data(), 1./d.dx, 1./d.dy,
                d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        r.save("strain2", s2);
    }

    template<Surface_model sm, int comp>
    void diff_uv(const Ref_run& r)
    {
        Ref_dims d(r); const size_t n = d.ncells, nk = d.kcells, n2 = d.ijcells;
        auto at = r.load("at", n); auto u = r.load("u", n); auto v = r.load("v", n); auto w = r.load("w", n);
        auto ev = r.load("evisc", n); auto fb = r.load("fluxbot", n2); auto ft = r.load("fluxtop", n2);
        auto dzi = r.load("dzi", nk); auto dzhi = r.load("dzhi", nk);
        auto rho = r.load("rhoref", nk); auto rhoh = r.load("rhorefh", nk);
        if (comp == 0)
            dk::diff_u<double, sm>(at.data(), u.data(), v.data(), w.data(), dzi.data(), dzhi.data(), 1./d.dx, 1./d.dy,
                    ev.data(), fb.data(), ft.data(), rho.data(), rhoh.data(), r.p("visc"),
                    d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        else
            dk::diff_v<double, sm>(at.data(), u.data(), v.data(), w.data(), dzi.data(), dzhi.data(), 1./d.dx, 1./d.dy,
                    ev.data(), fb.data(), ft.data(), rho.data(), rhoh.data(), r.p("visc"),
                    d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        r.save("at", at);
    }

    void diff_w(const Ref_run& r)
    {
        Ref_dims d(r); const size_t n = d.ncells, nk = d.kcells;
        auto at = r.load("at", n); auto u = r.load("u", n); auto v = r.load("v", n); auto w = r.load("w", n);
        auto ev = r.load("evisc", n);
        auto dzi = r.load("dzi", nk); auto dzhi = r.load("dzhi", nk);
        auto rho = r.load("rhoref", nk); auto rhoh = r.load("rhorefh", nk);
        dk::diff_w<double>(at.data(), u.data(), v.data(), w.data(), dzi.data(), dzhi.data(), 1./d.dx, 1./d.dy,
                ev.data(), rho.data(), rhoh.data(), r.p("visc"),
                d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        r.save("at", at);
    }

    template<Surface_model sm>
    void diff_c(const Ref_run& r)
    {
        Ref_dims d(r); const size_t n = d.ncells, nk = d.kcells, n2 = d.ijcells;
        auto at = r.load("at", n); auto a = r.load("a", n); auto ev = r.load("evisc", n);
        auto fb = r.load("fluxbot", n2); auto ft = r.load("fluxtop", n2);
        auto dzi = r.load("dzi", nk); auto dzhi = r.load("dzhi", nk);
        auto rho = r.load("rhoref", nk); auto rhoh = r.load("rhorefh", nk);
        dk::diff_c<double, sm>(at.data(), a.data(), dzi.data(), dzhi.data(), 1./(d.dx*d.dx), 1./(d.dy*d.dy),
                ev.data(), fb.data(), ft.data(), rho.data(), rhoh.data(), r.p("tPr"), r.p("visc"),
                d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        r.save("at", at);
    }

    void dnmul(const Ref_run& r)
    {
        Ref_dims d(r);
        auto ev = r.load("evisc", d.ncells); auto dzi = r.load("dzi", d.kcells);
        const double v = dk::calc_dnmul<double>(ev.data(), dzi.data(), 1./(d.dx*d.dx), 1./(d.dy*d.dy), r.p("tPr"),
                d.istart, d.iend, d.jstart, d.jend, d.kstart, d.kend, d.icells, d.ijcells);
        r.save("dnmul", std::vector<double>(1, v));
    }
}

void ref_register_diff()
{
    ref_register("strain2_nosfc", strain2<Surface_model::Disabled>);
    ref_register("strain2_sfc", strain2<Surface_model::Enabled>);
    ref_register("diff_u_nosfc", diff_uv<Surface_model::Disabled, 0>);
    ref_register("diff_u_sfc", diff_uv<Surface_model::Enabled, 0>);
    ref_register("diff_v_nosfc", diff_uv<Surface_model::Disabled, 1>);
    ref_register("diff_v_sfc", diff_uv<Surface_model::Enabled, 1>);
    ref_register("diff_w", diff_w);
    ref_register("diff_c_nosfc", diff_c<Surface_model::Disabled>);
    ref_register("diff_c_sfc", diff_c<Surface_model::Enabled>);
    ref_register("dnmul", dnmul);
}
