// pwh_refglobals.cpp -- definitions of the reference globals read by pwh_shim.cpp, with
// the names and types of global_variables.cpp:23-132, for programs (the tests) that
// link the shim without the reference's own global_variables.cpp.  A reference build
// links its global_variables.cpp instead of this file.
typedef double _Complex zcomplex;

int NR1 = 0, NR2 = 0, NR3 = 0, NNR = 0;
int NGWX = 0, NX = 0, NSP = 0, NAX = 0, NLMAX = 0, NKPT = 1;
int *NGW = 0, *NA = 0, *L_MAX = 0, *L_LOC = 0;
unsigned int *N123 = 0;
double *XKG = 0;
double TPIBA2 = 0.0;
double *WNL = 0;
double *PKG_A = 0;
zcomplex *EIGR = 0;
double OMEGA = 0.0;
double *WKPT = 0;
double *FOCC = 0;
double EKIN = 0.0;
zcomplex *C0 = 0;
bool USER_MAX_DIAG_ITER = false;
int MAX_DIAG_ITER = 0;
bool DIAG_VERBOSE = false;
int NEMPTY = 0;
int NG = 0;
int *N1 = 0, *N2 = 0, *N3 = 0;
double *G = 0, *VPS = 0, *RHOPS = 0;
zcomplex *SFAC = 0, *FNL = 0;
double ESR = 0.0, ESELF = 0.0, ETOT = 0.0, EHT = 0.0, EPSEU = 0.0, ENL = 0.0, EXC = 0.0;
double NEL = 0.0, KERKER_Q0 = 0.0;
int *MMAX = 0;
double *CLOG = 0, *R = 0, *PSI = 0, *VION = 0, *ZV = 0, *RGAUSS = 0, TPIBA = 0.0, *GGK = 0, *XK = 0, *PKG = 0;
