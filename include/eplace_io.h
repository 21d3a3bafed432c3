/*
 * eplace_io.h -- LEF/DEF design input and DEF output around the placer.
 *
 * Replaces the reference's DataBase load and store (vck5000/host/src/DataBase.cpp:7-96
 * readLEF/readDEF over the Limbo parser callbacks :318-424, and writeDEF :538-654).
 * The parser is this library's own: it reads the subset the reference's
 * callbacks use, plus what an ePlace netlist needs on top of it.
 *   LEF: MACRO name, SIZE w BY h, ORIGIN, PIN shapes (RECT / POLYGON); every
 *        other construct is skipped.
 *   DEF: DESIGN, UNITS DISTANCE MICRONS, DIEAREA, COMPONENTS (macro, PLACED /
 *        FIXED / COVER / UNPLACED, location, orientation), PINS (NET, DIRECTION,
 *        LAYER box, placement), NETS (component and IO pin connections);
 *        every other section and statement is skipped.
 * The netlist view (ep_netlist, include/eplace.h) is in microns: cells
 * [movable | fixed], movable = PLACED / UNPLACED components (or no status),
 * fixed = FIXED / COVER components followed by the IO pins (each a fixed cell
 * the size of its LAYER box).  A pin sits at the centre of its macro pin's
 * shapes, under the component's orientation; a macro pin without shapes (or an
 * unknown one) sits at the cell centre.  Nets with fewer than two pins carry
 * no wirelength and are left out of the view (they are written back).
 * Errors (EP_EIO: unreadable file or malformed statement; EP_EINVAL: a
 * component of an unknown macro, a net naming an unknown component or IO pin)
 * come back as status codes with ep_last_error() text; the reference asserts or
 * stops instead.
 */
#ifndef EPLACE_IO_H
#define EPLACE_IO_H

#include <stdint.h>

#include "eplace.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ep_lefdef ep_lefdef;

typedef struct ep_lefdef_info {
    int32_t units;            /* DEF database units per micron                    */
    int32_t n_macros;         /* LEF macros read                                  */
    int32_t n_components;     /* DEF components                                   */
    int32_t n_io_pins;        /* DEF PINS                                         */
    int32_t n_nets;           /* DEF NETS (all, including those left out of the view) */
    int32_t n_mov, n_fix;     /* netlist view: movable and fixed cells            */
    int32_t view_nets, view_pins;
    double die[4];            /* xl, yl, xh, yh in microns                        */
} ep_lefdef_info;

/* Reads n_lef LEF files (in order; a later macro of the same name replaces an
 * earlier one) and one DEF file. */
ep_status ep_lefdef_read(const char *const *lef_paths, int32_t n_lef, const char *def_path,
                         ep_lefdef **out);
/* DataBase(input_dir) (DataBase.cpp:7-96): every *.lef of the directory (by
 * name) and its floorplan.def; a directory without floorplan.def but with
 * exactly one .def file uses that one (the reference fails there). */
ep_status ep_lefdef_read_dir(const char *dir, ep_lefdef **out);
ep_status ep_lefdef_free(ep_lefdef *d);
/* Netlist view; its arrays belong to d and live until ep_lefdef_free. */
ep_status ep_lefdef_netlist(const ep_lefdef *d, ep_netlist *out);
/* design_name: at least 256 bytes, or NULL. */
ep_status ep_lefdef_get_info(const ep_lefdef *d, ep_lefdef_info *info, char *design_name);
/* DataBase::writeDEF (DataBase.cpp:538-654): writes <output_dir>/<design>.def
 * with the movable components at x, y (n_mov lower-left corners in microns, as
 * ep_get_positions returns them; the fillers after them are not read), rounded
 * to database units.  Fixed components keep their status and location
 * (the reference writes every component as PLACED). */
ep_status ep_lefdef_write_def(const ep_lefdef *d, const float *x, const float *y,
                              const char *output_dir);

#ifdef __cplusplus
}
#endif
#endif
