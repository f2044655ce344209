// photon_mi355x.h -- the reference-side binding of the MI355X photon engine.
//
// A new file for the reference tree (for example src/modules/photon_mi355x/),
// compiled with the reference's own hipcc line plus -I<this repository>/include
// and linked with libkpolrad.so.  It is a mesh module in place of
// particle::radiation::base_t (modules/particle/radiation/radiation.h:19-55
// over particle::base_t, modules/particle/driver.h:15-102): a problem
// generator enrols it with the reference's one-liner
//
//     mesh.enroll_module< photon_mi355x::base_t >(  );
//
// (mesh/mesh_enroll.h:77-109 registers its read / init / step closures;
// mesh_t::save calls its save).  Everything goes through the C ABI of
// include/kpolrad.h: no engine type crosses into the reference's code.
//
//   read  : the reference's [particle] keys (n_hv, n_cycle_stp, n_par,
//           step_max; radiation.h:26, driver.cpp:18, generate_par.h:37-41)
//           and the engine's [photon] keys (the same names as kp_module.h)
//   init  : the engine's block table straight from the reference mesh --
//           geo_t (geometry.h:22-40), region_logic_t, the face neighbours
//           exactly as proxy_base_t::setup reads them (proxy.h:22-62) --
//           on the device the reference device layer made current
//           (device/hip/hip.cpp), the per-cell medium from medium(), and
//           the tallies in memory from the reference device layer
//           (device.h:30-60, malloc_device)
//   step  : n_cycle_stp kp_forward calls of n_par packets per band with
//           fresh packet ids on the module's stream, then a stream sync,
//           as particle::base_t::step does (driver.cpp:55-60)
//   save  : J of each block under the tags rad_t::write gives it
//           (rad_data.cpp:45-49, dat_3d.h:117-126: block_<id_g>|rad_flx_
//           flux_{n_int,n_step,n_cell,n_gh,field}), FP64 [cell][band]
//
// One rank: a face neighbour on another rank (proxy.h:50-51) is rejected,
// since packets are not migrated between GPUs here.  Uniform block axes
// only (geo_t::dx0 > 0), cubic blocks of an even cell count.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kpolrad.h"
#include "mesh/mesh.h"

namespace photon_mi355x
{
// The reference's error behaviour: std::runtime_error (device/hip/hip.cpp:8-12).
inline void check( kp_status s )
{
    if( s != KP_OK )
        throw std::runtime_error( std::string( "photon_mi355x: " ) + kp_last_error(  ) );
}

////////////////////////////////////////////////////////////
// Block table of the engine from the reference mesh

struct table_t
{
    std::vector< double  > xf0, dx0, vol0;
    std::vector< int32_t > level, logic, nb_id;
    std::vector< uint8_t > nb_shift, phybnd;
    std::vector< int8_t  > d_lvl;
    std::vector< const mesh::block_t * > blocks; // by id_l
    kp_mesh_desc desc{};
};

// The local blocks of `mesh`, indexed by id_l, as a kp_mesh_desc whose
// arrays live in `t`.  Neighbour slot [a][side][j0][j1] and its shift bits
// follow proxy_base_t::setup (proxy.h:38-58): face neighbours (mode 0) only,
// side = the offset along a is positive, j0 / j1 = the sub-face offsets
// along (a+1)%3 / (a+2)%3, shift[i] from the reversed neighbour.
inline void describe( mesh::mesh_t & mesh, table_t & t )
{
    int nb = 0, n = -1;
    for( auto & b : mesh )
    {
        ++ nb;
        const auto & g = b.geo;
        if( g.n_ceff[ 0 ] != g.n_ceff[ 1 ] || g.n_ceff[ 0 ] != g.n_ceff[ 2 ] ||
            g.n_ceff[ 0 ] % 2 != 0 || ( n >= 0 && n != g.n_ceff[ 0 ] ) )
            throw std::runtime_error
                ( "photon_mi355x: needs cubic blocks of one even cell count" );
        n = g.n_ceff[ 0 ];
    }
    if( nb == 0 )
        throw std::runtime_error( "photon_mi355x: the mesh has no local block" );
    t.xf0.assign( 3 * nb, 0.0 );
    t.dx0.assign( 3 * nb, 0.0 );
    t.vol0.assign( nb, 0.0 );
    t.level.assign( nb, 0 );
    t.logic.assign( 3 * nb, 0 );
    t.nb_id.assign( 24 * nb, -1 );
    t.nb_shift.assign( 24 * nb, 0 );
    t.phybnd.assign( nb, 0 );
    t.d_lvl.assign( 6 * nb, 0 );
    t.blocks.assign( nb, nullptr );
    for( auto & b : mesh )
    {
        const int ib = b.id_l;
        if( ib < 0 || ib >= nb || t.blocks[ ib ] )
            throw std::runtime_error( "photon_mi355x: local ids are not 0..n-1" );
        t.blocks[ ib ] = & b;
        const auto & g = b.geo;
        for( int a = 0; a < 3; ++ a )
        {
            if( !( g.dx0[ a ] > 0 ) )
                throw std::runtime_error( "photon_mi355x: non-uniform block axis" );
            t.xf0  [ 3 * ib + a ] = g.xf0[ a ];
            t.dx0  [ 3 * ib + a ] = g.dx0[ a ];
            t.logic[ 3 * ib + a ] = b.reg[ a ];
            for( int s = 0; s < 2; ++ s )
                if( g.phybnd_flag[ a ][ s ] )
                    t.phybnd[ ib ] |= uint8_t( 1u << ( 2 * a + s ) );
        }
        t.vol0 [ ib ] = g.vol0 > 0 ? double( g.vol0 )
                      : t.dx0[ 3 * ib ] * t.dx0[ 3 * ib + 1 ] * t.dx0[ 3 * ib + 2 ];
        t.level[ ib ] = b.reg.level;
        for( const auto & [ nbr, reg ] : b.neighbors )
        {
            if( nbr.mode != 0 )
                continue;
            for( int a = 0; a < 3; ++ a )
            {
                if( nbr[ a ] >= 0 && nbr[ a ] <= 1 )
                    continue;
                if( nbr.guest_info.rank != b.rank )
                    throw std::runtime_error
                        ( "photon_mi355x: face neighbour on another rank" );
                const int side = nbr[ a ] > 0;
                const int j0 = nbr[ ( a + 1 ) % 3 ], j1 = nbr[ ( a + 2 ) % 3 ];
                const int slot = 24 * ib + ( ( a * 2 + side ) * 2 + j0 ) * 2 + j1;
                const auto nb_r = nbr.reverse(  );
                t.nb_id   [ slot ] = nbr.guest_info.id_l;
                t.nb_shift[ slot ] = uint8_t( ( nb_r[ ( a + 1 ) % 3 ] ? 1 : 0 ) |
                                              ( nb_r[ ( a + 2 ) % 3 ] ? 2 : 0 ) );
                t.d_lvl[ 6 * ib + 2 * a + side ] = int8_t( nbr.d_lvl );
            }
        }
    }
    auto & d = t.desc;
    d = kp_mesh_desc{};
    d.n_blocks = nb;
    d.n_cell   = n;
    const auto & lim = mesh.p_mgn->x_lim_global;
    for( int a = 0; a < 3; ++ a )
    {
        d.periodic[ a ] = mesh.tree(  ).periodic[ a ] ? 1 : 0;
        d.x_min   [ a ] = lim[ 0 ][ a ];
        d.x_max   [ a ] = lim[ 1 ][ a ];
    }
    d.xf0      = t.xf0.data(  );
    d.dx0      = t.dx0.data(  );
    d.vol0     = t.vol0.data(  );
    d.level    = t.level.data(  );
    d.logic    = t.logic.data(  );
    d.nb_id    = t.nb_id.data(  );
    d.nb_shift = t.nb_shift.data(  );
    d.d_lvl    = t.d_lvl.data(  );
    d.phybnd   = t.phybnd.data(  );
    d.coord    = KP_COORD_CARTESIAN;
    d.xf       = nullptr;
    d.sph      = nullptr;
}

////////////////////////////////////////////////////////////
// The module

class base_t : public mesh::mod_base_t
{
public:
    ~base_t(  )
    {
        if( eng_ )
            kp_engine_destroy( eng_ );
    }

    // Medium of cell (i0,i1,i2) of block b: extinction kappa*rho per unit
    // length (the rad_t::mfp_i field, rad_data.h:16), albedo and HG g.
    // Default: the uniform [photon] kappa / albedo / g; a problem generator
    // overrides it like init_cond.
    virtual void medium( const mesh::block_t & b, int i0, int i1, int i2,
                         float & kappa_rho, float & albedo, float & g ) const
    {
        ( void ) b; ( void ) i0; ( void ) i1; ( void ) i2;
        kappa_rho = kappa_;
        albedo    = albedo_;
        g         = g_;
    }

    void read( const input & args ) override
    {
        n_hv_   = args.get< int    >( "particle", "n_hv",        1   );
        n_cyc_  = args.get< int    >( "particle", "n_cycle_stp", 1   );
        n_par_  = args.get< double >( "particle", "n_par",       1.0 );
        steps_  = args.get< int    >( "particle", "step_max",    100 );
        seed_   = args.get< int    >( "photon",   "seed",        42  );
        const auto model = args.get< std::string >( "photon", "model", "thomson" );
        if( model != "thomson" && model != "hg_rayleigh" )
            throw std::runtime_error( "photon_mi355x: photon|model must be thomson or hg_rayleigh" );
        model_  = model == "thomson" ? KP_THOMSON : KP_HG_RAYLEIGH;
        p_max_  = args.get< double >( "photon", "p_max",  1.0 );
        kappa_  = args.get< float  >( "photon", "kappa",  0.0f );
        albedo_ = args.get< float  >( "photon", "albedo", 1.0f );
        g_      = args.get< float  >( "photon", "g",      0.0f );
        weight_ = args.get< double >( "photon", "weight", 1.0 );
        image_n_    = args.get< int    >( "photon", "image_n",    128 );
        image_half_ = args.get< double >( "photon", "image_half", 0.5 );
        rr_weight_  = args.get< double >( "photon", "rr_weight",  0.01 );
        rr_survive_ = args.get< double >( "photon", "rr_survive", 0.1 );
        precision_  = args.get< std::string >( "photon", "precision", "f64" ) == "f32"
                    ? KP_F32 : KP_F64;
        kappa_scale_.assign( n_hv_, 1.0 );
        std::vector< double > ks;
        if( args( "photon", "kappa_scale", ks ) )
        {
            if( int( ks.size(  ) ) != n_hv_ )
                throw std::runtime_error( "photon_mi355x: photon|kappa_scale needs n_hv values" );
            kappa_scale_ = ks;
        }
        incl_.clear(  );
        args( "photon", "incl", incl_ );
        has_source_ = args( "photon", "source", source_ );
    }

    void init( const input & args, mesh::mesh_t & mesh ) override
    {
        ( void ) args;
        p_dev = mesh.p_dev;
        // the reference device layer made the rank's device current
        int device = 0;
        if( hipGetDevice( & device ) != hipSuccess )
            throw std::runtime_error( "photon_mi355x: no current HIP device" );
        describe( mesh, table_ );
        check( kp_engine_create( & table_.desc, device, & eng_ ) );
        const int n  = table_.desc.n_cell;
        const size_t n3 = size_t( n ) * n * n;
        n_cells_ = n3 * table_.blocks.size(  );
        std::vector< float > kap( n_cells_ ), alb( n_cells_ ), gg( n_cells_ );
        for( size_t ib = 0; ib < table_.blocks.size(  ); ++ ib )
            for( int i2 = 0; i2 < n; ++ i2 )
                for( int i1 = 0; i1 < n; ++ i1 )
                    for( int i0 = 0; i0 < n; ++ i0 )
                    {   // dat_3d_t order (dat_3d.h:145-147), no ghosts
                        const size_t c = ib * n3 + i0 + size_t( n ) * ( i1 + size_t( n ) * i2 );
                        medium( * table_.blocks[ ib ], i0, i1, i2, kap[ c ], alb[ c ], gg[ c ] );
                    }
        check( kp_set_media( eng_, kap.data(  ), alb.data(  ), gg.data(  ) ) );
        if( ! has_source_ )
            for( int a = 0; a < 3; ++ a )
                source_[ a ] = 0.5 * ( table_.desc.x_min[ a ] + table_.desc.x_max[ a ] );
        stream_ = mesh.p_dev->yield_stream(  );
        n_img_  = size_t( n_hv_ ) * incl_.size(  ) * image_n_ * image_n_ * 3;
        J_   = p_dev->malloc_device< double   >( n_cells_ * n_hv_ );
        img_ = p_dev->malloc_device< double   >( n_img_ > 0 ? n_img_ : 1 );
        cnt_ = p_dev->malloc_device< uint64_t >( KP_NCOUNTERS );
        if( hipMemsetAsync( J_, 0, n_cells_ * n_hv_ * sizeof( double ), stream_ ) != hipSuccess ||
            hipMemsetAsync( img_, 0, ( n_img_ > 0 ? n_img_ : 1 ) * sizeof( double ), stream_ ) != hipSuccess ||
            hipMemsetAsync( cnt_, 0, KP_NCOUNTERS * sizeof( uint64_t ), stream_ ) != hipSuccess )
            throw std::runtime_error( "photon_mi355x: clearing the tallies failed" );
        p_dev->sync_stream( stream_ );
    }

    // particle::base_t::step (driver.cpp:55-60): n_cycle_stp integrate
    // cycles, then the stream sync.
    void step( mesh::mesh_t & mesh ) override
    {
        ( void ) mesh;
        std::vector< kp_band > bands( n_hv_ );
        for( int b = 0; b < n_hv_; ++ b )
        {
            bands[ b ] = kp_band{  };
            bands[ b ].kappa_scale = kappa_scale_[ b ];
            bands[ b ].albedo      = 1.0;
            bands[ b ].g           = 0.0;
            bands[ b ].n_packets   = uint64_t( n_par_ );
            bands[ b ].flags       = KP_BAND_CELL_ALBEDO | KP_BAND_CELL_G;
        }
        kp_run run{  };
        run.seed          = uint32_t( seed_ );
        run.scatter_model = model_;
        run.p_max         = p_max_;
        run.n_bands       = n_hv_;
        run.precision     = precision_;
        run.bands         = bands.data(  );
        run.step_max      = steps_;
        run.n_incl        = int32_t( incl_.size(  ) );
        run.incl_deg      = incl_.empty(  ) ? nullptr : incl_.data(  );
        run.rr_weight     = rr_weight_;
        run.rr_survive    = rr_survive_;
        run.image_n       = image_n_;
        run.image_half    = image_half_;
        kp_source src{  };
        for( int a = 0; a < 3; ++ a )
            src.x[ a ] = source_[ a ];
        src.weight = weight_;
        kp_tallies out{ J_, n_img_ > 0 ? img_ : nullptr, cnt_, nullptr };
        for( int c = 0; c < n_cyc_; ++ c )
        {
            run.packet_offset = next_id_;
            check( kp_forward( eng_, & src, & run, & out, ( void * ) stream_ ) );
            next_id_ += uint64_t( n_par_ );
        }
        p_dev->sync_stream( stream_ );
    }

    // rad_t::write (rad_data.cpp:45-49): block b's slice of J, n_cell^3
    // cells x n_hv bands, is exactly its flx field (n_int = n_hv, no ghosts).
    void save( mesh::mesh_t & mesh, binary_io::base_t & bio ) override
    {
        ( void ) mesh;
        const int n = table_.desc.n_cell;
        const size_t per = size_t( n ) * n * n * n_hv_;
        std::vector< double > host( per * table_.blocks.size(  ) );
        p_dev->cp( host.data(  ), J_, host.size(  ) );
        const int n_step = 0, n_cell[ 3 ] = { n, n, n }, n_gh[ 3 ] = { 0, 0, 0 };
        for( size_t ib = 0; ib < table_.blocks.size(  ); ++ ib )
        {
            const std::string pre = table_.blocks[ ib ]->io_prefix(  ) + "rad_flx_flux_";
            bio.write( & n_hv_,  1, sizeof( int ), pre + "n_int",  true );
            bio.write( & n_step, 1, sizeof( int ), pre + "n_step", true );
            bio.write( n_cell,   3, sizeof( int ), pre + "n_cell", true );
            bio.write( n_gh,     3, sizeof( int ), pre + "n_gh",   true );
            bio.write( host.data(  ) + ib * per, per, sizeof( double ), pre + "field", true );
        }
    }

    void finalize( mesh::mesh_t & mesh ) override
    {
        ( void ) mesh;
        if( J_ )   p_dev->free_device( J_ );
        if( img_ ) p_dev->free_device( img_ );
        if( cnt_ ) p_dev->free_device( cnt_ );
        if( eng_ )
            kp_engine_destroy( eng_ );
        eng_ = nullptr;
    }

    int space_order(  ) override { return 1; }

    const table_t & table(  ) const { return table_; }

protected:
    table_t                 table_;
    kp_engine             * eng_ = nullptr;
    device::stream_t        stream_{  };
    double                * J_   = nullptr;
    double                * img_ = nullptr;
    uint64_t              * cnt_ = nullptr;
    size_t                  n_cells_ = 0, n_img_ = 0;
    uint64_t                next_id_ = 0;
    int                     n_hv_ = 1, n_cyc_ = 1, steps_ = 100, seed_ = 42;
    int                     model_ = KP_THOMSON, precision_ = KP_F64;
    int                     image_n_ = 128;
    double                  n_par_ = 1.0, p_max_ = 1.0, weight_ = 1.0;
    double                  image_half_ = 0.5, rr_weight_ = 0.01, rr_survive_ = 0.1;
    float                   kappa_ = 0.0f, albedo_ = 1.0f, g_ = 0.0f;
    std::vector< double >   kappa_scale_, incl_;
    std::array< double, 3 > source_{ { 0.0, 0.0, 0.0 } };
    bool                    has_source_ = false;
};
};                              // namespace photon_mi355x
