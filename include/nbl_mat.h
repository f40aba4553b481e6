/*
 * nbl_mat.h -- C ABI of libnbl_mat.so: the reference's HDF5 material files
 * (cstool output, read by io/hdf5_file.cpp and the physics create()
 * functions) in and out of the raw-table structs of nbl_hip.h.
 *
 * A separate library from libnbl_hip.so because it links HDF5 (1.10, C and
 * high-level APIs); libnbl_hip.so itself depends only on the HIP runtime.
 * Every function returns NBL_OK or an NBL_ERR_* code and leaves the message
 * for nbl_mat_last_error().
 *
 * File layout (what the reference reads):
 *   root attributes      name, cstool_version: strings;
 *                        barrier, fermi, band_gap (optional, -1 eV if absent),
 *                        phonon_loss, effective_A: compound {value: f64, unit: string}
 *   /kieft/elastic/imfp               [K]     (length^-1)
 *   /kieft/elastic/costheta_icdf      [K][P]
 *   /kieft/inelastic/imfp             [K]     (length^-1)
 *   /kieft/inelastic/w0_icdf          [K][P]  (energy)
 *   /kieft/ionization/binding_icdf    [K][P]  (energy)
 *   /kieft/ionization/outer_shells    [n]     (energy)
 *   /full_penn/imfp                   [K]     (length^-1)
 *   /full_penn/omega_icdf             [K][P]  (energy)
 *   /full_penn/q_icdf                 [K][omega/K][P]   (length^-1)
 *   /ionization/binding_icdf          [K][omega/K][P]   (energy)
 * Each table dataset has a string attribute "units"; its axes are HDF5
 * dimension scales (each with a "units" attribute), log-spaced for energies
 * and the binding table's omega/K, linear otherwise; a linear axis without a
 * scale is 0..1 (hdf5_file.cpp:291-326).
 */
#ifndef NBL_MAT_H
#define NBL_MAT_H

#include "nbl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nbl_matfile_s* nbl_matfile;

/* Message of the last failed call on this thread (empty if none). */
const char* nbl_mat_last_error(void);

/* hdf5_file::hdf5_file (io/hdf5_file.cpp:157-164): opens read-only. */
int nbl_mat_open(const char* path, nbl_matfile* out);
int nbl_mat_close(nbl_matfile f);

/* hdf5_file::exists (io/hdf5_file.cpp:175-178): *out = 1 if the path names
 * a group or dataset. */
int nbl_mat_exists(nbl_matfile f, const char* path, int32_t* out);

/* hdf5_file::get_property_string (io/hdf5_file.cpp:180-189): the string
 * attribute `name` of the root; *len = its length, up to cap - 1 bytes and a
 * NUL copied to buf (buf may be NULL with cap 0 to ask for the length). */
int nbl_mat_get_string(nbl_matfile f, const char* name, char* buf, uint64_t cap, uint64_t* len);

/* hdf5_file::get_property_quantity (io/hdf5_file.cpp:190-201): value in the
 * reference's unit system (eV, nm, ns, K, e: units/unit_system.h:16-29) and
 * the powers of (energy, length, time, temperature, charge). */
int nbl_mat_get_quantity(nbl_matfile f, const char* name, double* value, int32_t dims[5]);

/* unit_parser::parse_unit with the default parser (units/unit_parser.h:46-92,
 * units/unit_system.h:33-47): "eV", "nm^-1", "g/cm^3", ""; unknown names fail. */
int nbl_mat_parse_unit(const char* text, double* value, int32_t dims[5]);

/* The Kieft-Bosch raw tables as kieft_elastic::create / kieft_inelastic::create
 * and material::create read them (physics/kieft/elastic.h:120-166,
 * physics/kieft/inelastic.h:212-327, core/material.inl:25-34), in eV and
 * nm^-1: table values are multiplied by their unit in float as the reference
 * does before its post-processing.  The five tables must share one energy
 * and one probability axis.  Arrays are owned by f (valid until close). */
int nbl_mat_kieft_raw(nbl_matfile f, nbl_kieft_raw* out);

/* The full Penn raw tables as full_penn::create reads them
 * (physics/full_penn.h:294-399); imfp and omega_icdf share their energy axis. */
int nbl_mat_penn_raw(nbl_matfile f, nbl_penn_raw* out);

/* Writes a material file in the layout above (the Kieft groups; the full
 * Penn groups too when penn != NULL; penn's fermi and band gap must equal
 * kieft's).  Units: eV, nm^-1, g. */
int nbl_mat_write(const char* path, const char* name, const nbl_kieft_raw* kieft, const nbl_penn_raw* penn);

#ifdef __cplusplus
}
#endif
#endif
