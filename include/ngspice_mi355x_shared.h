/*
 * ngspice_mi355x_shared.h -- the embedding interface of ngspice's shared
 * library (src/include/ngspice/sharedspice.h:176-220, 247-319, 384-486), exported by
 * the MI355X engine library for BSIM4 transient decks.
 *
 * The types and signatures are those of sharedspice.h, field for field, so a
 * program written against the reference's header links against this
 * library unchanged.  What runs behind them is the engine: the deck goes
 * through the engine's front end (ngmi_netlist_parse), the analysis through
 * the batched GPU transient (one circuit), and SendData receives every
 * accepted point.
 *
 * Commands ngSpice_Command understands (a subset of the reference's
 * front-end commands, sharedspice.c:1124-1159 / runc):
 *   source <file>        read a deck from a file
 *   circbyline <line>    add one deck line; ".end" completes the deck
 *   run                  run the deck's .tran (foreground)
 *   bg_run               the same on a background thread (BGThreadRunning
 *                        reports start and end)
 *   tran <tstep> <tstop> [<tstart> [<tmax>]]   run this transient instead
 *   remcirc / reset      drop the circuit; destroy [all] drops the results
 *   echo <text>          print through SendChar
 *   quit / exit          ControlledExit(0, false, true, ...)
 * Anything else returns 1 and prints "stderr Error: ..." through SendChar.
 */
#ifndef NGSPICE_MI355X_SHARED_H
#define NGSPICE_MI355X_SHARED_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ngspice_NGSPICE_H
struct ngcomplex {
    double cx_real;
    double cx_imag;
};
typedef struct ngcomplex ngcomplex_t;
#endif

typedef bool NG_BOOL;

/* sharedspice.h:176-183 */
typedef struct vector_info {
    char *v_name;
    int v_type;
    short v_flags;
    double *v_realdata;
    ngcomplex_t *v_compdata;
    int v_length;
} vector_info, *pvector_info;

/* sharedspice.h:185-198 */
typedef struct vecvalues {
    char *name;
    double creal;
    double cimag;
    NG_BOOL is_scale;
    NG_BOOL is_complex;
} vecvalues, *pvecvalues;

typedef struct vecvaluesall {
    int veccount;
    int vecindex;
    pvecvalues *vecsa;
} vecvaluesall, *pvecvaluesall;

/* sharedspice.h:200-220 */
typedef struct vecinfo {
    int number;
    char *vecname;
    NG_BOOL is_real;
    void *pdvec;
    void *pdvecscale;
} vecinfo, *pvecinfo;

typedef struct vecinfoall {
    char *name;
    char *title;
    char *date;
    char *type;
    int veccount;
    pvecinfo *vecs;
} vecinfoall, *pvecinfoall;

/* callbacks, sharedspice.h:247-319 */
typedef int(SendChar)(char *, int, void *);
typedef int(SendStat)(char *, int, void *);
typedef int(ControlledExit)(int, NG_BOOL, NG_BOOL, int, void *);
typedef int(SendData)(pvecvaluesall, int, int, void *);
typedef int(SendInitData)(pvecinfoall, int, void *);
typedef int(BGThreadRunning)(NG_BOOL, int, void *);

/* ngSpice_Init (sharedspice.h:384-385): stores the callbacks (any may be
 * NULL) and the user pointer; returns 0. */
int ngSpice_Init(SendChar *printfcn, SendStat *statfcn, ControlledExit *ngexit, SendData *sdata,
                 SendInitData *sinitdata, BGThreadRunning *bgtrun, void *userData);
/* ngSpice_Command (sharedspice.h:403): 0 on success, 1 on error or before
 * ngSpice_Init; NULL clears the deck being collected by circbyline. */
int ngSpice_Command(char *command);
/* ngSpice_Circ (sharedspice.h:465): the deck as lines, a ".end" line, then
 * NULL; 0 on success, 1 when the deck does not parse. */
int ngSpice_Circ(char **circarray);
/* ngGet_Vec_Info (sharedspice.h:407): a vector of the last run ("time",
 * a node name or v(name), a source's "<name>#branch" or i(name)), case
 * insensitive; NULL when absent.  Valid until the next run or destroy. */
pvector_info ngGet_Vec_Info(char *vecname);
/* ngSpice_CurPlot / ngSpice_AllPlots / ngSpice_AllVecs (sharedspice.h:
 * 470-482): "tran1" after a run ("const" before), NULL-terminated arrays */
char *ngSpice_CurPlot(void);
char **ngSpice_AllPlots(void);
char **ngSpice_AllVecs(char *plotname);
/* ngSpice_running (sharedspice.h:486): true while a bg_run is active */
NG_BOOL ngSpice_running(void);

#ifdef __cplusplus
}
#endif
#endif
