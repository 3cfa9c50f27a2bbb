/* Test infrastructure only (the checker, never the product).
 *
 * CPU restatement of one rank's CABARET/ARK time step of the reference
 * solver (source/main.cpp:80-104 and the phase functions it calls), for a
 * triply periodic box in which the reference's MPI halo exchange
 * (source/main.cpp:603-801, 1049-1631) reduces to a periodic ghost fill.
 *
 * Layout is the reference's: every field is a full x-fastest array with the
 * reference's extents (source/main.cpp:198-277; index map source/Arr3d.cpp:19-21),
 * cells (N1+2)(N2+2)(N3+2), faces and stresses (N1+3)(N2+3)(N3+3), where
 * Nd is the number of interior cells along axis d (the reference's nd-1).
 */
#ifndef ARK_ORACLE_H
#define ARK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Field slots, in the order used by every tool in the project. */
enum {
    ORC_roCon, ORC_u1Con, ORC_u2Con, ORC_u3Con, ORC_tCon,
    ORC_ronCon, ORC_u1nCon, ORC_u2nCon, ORC_u3nCon, ORC_tnCon,
    ORC_f1, ORC_f2, ORC_f3,
    ORC_condition,
    ORC_ro1, ORC_t1, ORC_u11, ORC_u21, ORC_u31, ORC_p1,
    ORC_ro2, ORC_t2, ORC_u12, ORC_u22, ORC_u32, ORC_p2,
    ORC_ro3, ORC_t3, ORC_u13, ORC_u23, ORC_u33, ORC_p3,
    ORC_sigm11, ORC_sigm21, ORC_sigm31,
    ORC_sigm12, ORC_sigm22, ORC_sigm32,
    ORC_sigm13, ORC_sigm23, ORC_sigm33,
    ORC_NUM_FIELDS
};

typedef struct {
    int N1, N2, N3;
    double dx1, dx2, dx3;
    double dt, CFL, VIS, sound, ro0, t0;
    double* f[ORC_NUM_FIELDS];
    int scheme;   /* 0: the reference (gq from R on X1/X3); 1: gq from Q on every axis */
    int boundary; /* 0: periodic box; 1: the reference's channel (see ark_oracle.c) */
} ark_oracle_state;

/* The reference's own problem (Input/InitializeData, source/main.cpp:118-372)
 * on one rank with the given node counts and lengths: level-n fields, interior
 * faces and the solid mask; every other element 0; Con := nCon. */
void ark_oracle_init_channel(ark_oracle_state* s, int n1_g, int n2_g, int n3_g, double x1_e, double x2_n,
                             double x3_t, double u10);

/* 1 for face/stress-sized fields, 0 for cell-sized ones. */
int ark_oracle_is_face(int slot);

/* Each function mirrors the reference function of the same name. */
void ark_oracle_time_step_size(ark_oracle_state* s);   /* sets s->dt */
void ark_oracle_phase1(ark_oracle_state* s);
void ark_oracle_stress_tensor(ark_oracle_state* s);
void ark_oracle_use_forces(ark_oracle_state* s);
void ark_oracle_phase2(ark_oracle_state* s);
/* nsteps full steps in the reference order; dt_out (may be NULL) receives each step's dt. */
void ark_oracle_step(ark_oracle_state* s, int nsteps, double* dt_out);
/* Taylor-Green vortex initial state of the oracle harness (ref_harness.cpp
 * init_tgv): u=U0 sin x cos y cos z, v=-U0 cos x sin y cos z, w=0,
 * p=(rho0 U0^2/16)(cos2x+cos2y)(cos2z+2), rho=rho0+p/c^2, t=t0, evaluated at
 * the periodic image of every point (ghosts are bitwise copies), Con := nCon,
 * f and sigma zero, condition zero.  Sets every array in full. */
void ark_oracle_init_tgv(ark_oracle_state* s, double U0);
/* Reference kinetic-energy definition (WriteEnergy) and sum of rho*dvc, interior cells of nCon. */
double ark_oracle_energy(const ark_oracle_state* s);
double ark_oracle_mass(const ark_oracle_state* s);

#ifdef __cplusplus
}
#endif
#endif
