// enroll_example.cpp -- a problem generator of the reference that runs the
// MI355X photon engine (prob::run, user/probgen.cpp:18-22, overridden as the
// reference intends).  The same lines a maintainer writes to switch a
// problem from particle::radiation::base_t to this engine; it also serves as
// the translation unit that checks photon_mi355x.h against the reference's
// headers (hipcc -fsyntax-only with -I<reference>/src).
#include "mesh/mesh.h"
#include "mesh/mesh_enroll.h"
#include "user/probgen.h"

#include "photon_mi355x.h"

namespace
{
// A medium the problem defines cell by cell, like init_cond does for
// block data: a uniform sphere of kappa*rho = 2.5 and radius 0.4 around
// the domain centre.
struct sphere_t : public photon_mi355x::base_t
{
    void medium( const mesh::block_t & b, int i0, int i1, int i2,
                 float & kappa_rho, float & albedo, float & g ) const override
    {
        const int i[ 3 ] = { i0, i1, i2 };
        double r2 = 0;
        for( int a = 0; a < 3; ++ a )
        {
            const double x = b.geo.x_cc( a, i[ a ] ) - 0.5;
            r2 += x * x;
        }
        kappa_rho = r2 < 0.16 ? 2.5f : 0.0f;
        albedo    = 1.0f;
        g         = 0.0f;
    }
};
}   // namespace

namespace prob
{
void run( int argc, char * argv[  ] )
{
    input args;
    cmd( argc, argv, args );
    mesh::mesh_t mesh;
    auto photon = mesh.enroll_module< sphere_t >(  );
    mesh.init  ( args );
    mesh.evolve(      );
    ( void ) photon;
}
};                              // namespace prob
