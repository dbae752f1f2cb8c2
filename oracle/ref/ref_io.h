/*
 * ref_io.h -- file protocol of the golden-vector generator (oracle/_ref).
 *
 * TEST INFRASTRUCTURE ONLY.  The generator wraps the reference's own
 * translation units, compiled from /root/reference by oracle/Makefile; this
 * header holds only the project's I/O glue: a run directory with params.txt
 * ("key value" lines) and raw float64 arrays <name>.bin; outputs are written
 * as out_<name>.bin.
 */
#ifndef MHH_REF_IO_H
#define MHH_REF_IO_H
#include <map>
#include <string>
#include <vector>

struct Ref_run
{
    std::string dir;
    std::map<std::string, double> par;
    double p(const std::string& k) const;
    int pi(const std::string& k) const { return static_cast<int>(p(k)); }
    std::vector<double> load(const std::string& name, size_t n) const;
    void save(const std::string& name, const std::vector<double>& a) const;
};

/* Bounds and strides of the reference kernels. */
struct Ref_dims
{
    int istart, iend, jstart, jend, kstart, kend, icells, jcells, ijcells, ncells, kcells;
    int itot, jtot, ktot;
    double dx, dy;
    explicit Ref_dims(const Ref_run& r);
};

typedef void (*ref_op)(const Ref_run&);
void ref_register(const char* name, ref_op op);
#endif
