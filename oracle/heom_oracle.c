/*
 * heom_oracle.c -- CPU oracle (FP64) for DM-HEOM's population-dynamics and
 * linear-absorption path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker.  Only tests/, the smoke check
 * and the benchmark's CPU-baseline leg load it; the product never links or calls it.
 *
 * Each function restates the reference's algorithm statement by statement, in FP64,
 * compiled with -ffp-contract=off so that every statement rounds as written.  Paths
 * are relative to the reference repository root (dm-heom/...).
 *
 * Parity status of each part (how it is pinned is in the tests):
 *   partitions / graph   pinned: closed-form counts, brute-force enumeration, a step-by-step
 *                        restatement of the reference generator, plus/minus invariants
 *   precompute           pinned: 50-digit mpmath evaluation of the formulas; Pade table
 *                        checked against the Bose function it approximates; c_k, nu_k,
 *                        Delta_b and Phi_b against the Drude-Lorentz correlation function
 *                        computed by Fourier quadrature of J(w) (the exponentials converge
 *                        to Re C(t) with every Pade term, Delta_b is the zero-frequency
 *                        weight they miss, Phi_b gives Im C(t)); planted factors fail
 *   rhs                  pinned: dense-Liouvillian brute force, no-bath von Neumann limit
 *                        (expm), trace conservation, Hermiticity
 *   rk4                  pinned: order-4 convergence against expm of the dense Liouvillian;
 *                        the exact call sequence of the un-vendored noma/num stepper is
 *                        parity unpinned (classical tableau assumed)
 *   norms, mmult_left,   pinned: textbook definitions checked in numpy
 *   mmult_right, trace
 *   propagate_timedep    pinned: a zero field is plain RK4 bit for bit; with a pulse, step
 *                        halving converges at fourth order to a fine-step reference (fails
 *                        if a stage sees the Hamiltonian at the wrong time)
 *   rhs_tss, tss_step    pinned: f_TSS equals eps y plus the ordinary RHS with H = 0 and
 *                        S_u = 0; the converged search equals the null vector of the dense
 *                        Liouvillian (SVD), and a planted sign error in the denominator
 *                        converges elsewhere; the Boltzmann start against scipy's expm
 *   adjoint              no reference counterpart; pinned: <mu, L y> = <L^dag mu, y> on random
 *                        states against the pinned rhs, every coefficient gradient against
 *                        central differences of oracle_propagate, and the discrete-adjoint
 *                        gradient against central differences of whole trajectories
 */
#include "heom_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ħ as used by the kernels (cl/heom/common.cl:45) and the host
 * (include/heom/constants.hpp:22); k_B (include/heom/constants.hpp:23). */
static const double ORACLE_HBAR = 1.05457148e-34;
static const double ORACLE_KB = 1.380650239684840458953781387048428305608616777679979316e-23;

/* Bosonic Pade expansion tables, rows of 10, row K-1 used for K "matsubaras"
 * (src/heom/constants.cpp:13-76).  These are data: reproduced value for value. */
#define ORACLE_MAX_PADE 10
static const long double oracle_eta_pade[100] = {
    2.5L, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    5.967175818975844784692668313888340351402L, 1.032824181024155215307331686111659648598L,
    0, 0, 0, 0, 0, 0, 0, 0,
    11.19885869716451815284926959121234309769L, 1.300914314466752848423178399066599315543L,
    1.000226988368728998727552009721057586765L, 0, 0, 0, 0, 0, 0, 0,
    18.0790807744036381353279225690526377613L, 1.905605223763320862837482220696034638651L,
    1.015313588078062122619093914564704203251L, 1.000000413754978879215501295686623396803L,
    0, 0, 0, 0, 0, 0,
    26.58655772186703141593168716080327785988L, 2.800146619916205478116315260146686752654L,
    1.11303340660549395961527574088616034253L, 1.000262251326821823459883045067529050956L,
    1.000000000284447322876838793096345993981L, 0, 0, 0, 0, 0,
    36.71706591275454016953866376131015396352L, 3.915524608369778138857104011713377210806L,
    1.36060762304475364762294649920250800437L, 1.006800099494035677225303524345587578974L,
    1.000001756336801785832037797152047562438L, 1.000000000000090580923944406276325679894L,
    0, 0, 0, 0,
    48.46945901876815939911017902590799133678L, 5.221358700226750582830885303574125272024L,
    1.76101413086885578701660606207054290627L, 1.048001360645944284478966960182497444605L,
    1.000166783806547433957971903891016917251L, 1.000000005683742497388628469835034317229L,
    1.000000000000000015216762274538791805838L, 0, 0, 0,
    61.84333899016684164883279780411421181762L, 6.709368642028293738757159220036413607145L,
    2.276634739389544433995764512332145784496L, 1.167662545543253700136635145976136601895L,
    1.002993112803799185748365319635086694072L, 1.000001970058157018499040773410845478966L,
    1.000000000010110274028760159309306116011L, 1.000000000000000000001477065185853899797L,
    0, 0,
    76.83854093567190329220602502323513693626L, 8.377617889976198321037045566932019912232L,
    2.875743265352905559741498484957445572073L, 1.386916450076667643674849716161379189731L,
    1.021092605957815260732482889860564891757L, 1.000088840018504469980284336164406834212L,
    1.000000012945994679324047582718370173438L, 1.000000000000010773303766311318241171588L,
    1.000000000000000000000000088652435318708L, 0,
    93.45498712043917746792411329602543487242L, 10.22570574466955029693172438665880309108L,
    3.544527893692194584758656451673749774362L, 1.692413746219203914341865930002582507947L,
    1.081051704862394372162728930732659549555L, 1.00131233986343205708864037801373384889L,
    1.000001450202211357298203645556646479551L, 1.000000000051835942172273790638144172787L,
    1.000000000000000007321793190694778206164L, 1.000000000000000000000000000003467497243L};

static const long double oracle_xi_pade[100] = {
    7.745966692414833770358530799564799221666L, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    19.49961875292266554276226912855528560143L, 6.305939144224808474630860431551181929313L,
    0, 0, 0, 0, 0, 0, 0, 0,
    36.11928942365857887916363930568047443027L, 12.95828669566379113836433196980720441554L,
    6.283290335721270010976001332769855412497L, 0, 0, 0, 0, 0, 0, 0,
    57.78794000633463555484777318092706162218L, 20.56259756752951982974170033395653675754L,
    12.57995038432843325319503152817888327099L, 6.283185452295109512774977729104135871222L,
    0, 0, 0, 0, 0, 0,
    84.53692559026397942121263237089439381387L, 29.57927623718706979609581199085607351807L,
    19.00468956251068382661480700055507075954L, 12.56654232705146493528755058935472519371L,
    6.28318530726039477349210954220860566044L, 0, 0, 0, 0, 0,
    116.3739347608330350073189696097089226231L, 40.21293928856715127416984659035932828755L,
    25.82783232953028238453829864275777795836L, 18.8562400267953494905521831656114592347L,
    12.56637153917480629098401692212665860923L, 6.28318530717960815500107855468522104325L,
    0, 0, 0, 0,
    153.3015628393653726331899655194706876159L, 52.52412157944826796685914142534634337648L,
    33.34898703451500336730287935111300571215L, 25.19900528903505310331751346668259596763L,
    18.84968412056844144448843441419311682201L, 12.56637061688110415733890859371956448695L,
    6.283185307179586480075153567585578254575L, 0, 0, 0,
    195.3208905359343070977627172198676228279L, 66.52807871550278259674646939491255055216L,
    41.76755706259573674247409323452395574559L, 31.72895561746715247842724612158554657776L,
    25.13585329999435832854980649807171916496L, 18.84955718260564459859203580283272390792L,
    12.56637061436306356083304954158209077981L, 6.283185307179586476925556323382874292731L,
    0, 0,
    242.43243633332578223865362314113358865L, 82.22872388545609111729841286466997903147L,
    51.17127400682872814625269028495752604681L, 38.64346654893615544641179566060655896253L,
    31.44483562147124840697966631064127684252L, 25.13281634168110515436812140082770700199L,
    18.84955592869136612218434285960431278578L, 12.56637061435917662145272429955102313797L,
    6.283185307179586476925286781031145512608L, 0,
    294.6364754696112018848669150774790893224L, 99.62709625387022543419284510745749131322L,
    61.58845079917599582892471747985579618666L, 46.1203736291541763005458261878858666512L,
    37.84390364517245088090747934823213310617L, 31.41734004510526384373874233227272273452L,
    25.13274227487208277473085335076650873783L, 18.84955592156405141516275123243075673053L,
    12.56637061435917295608825501073938367082L, 6.283185307179586476925286766559517951084L};

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================== */
/* integer partitions (src/heom/integer_partition.cpp)                       */
/* ======================================================================== */

/* partitions of n into exactly k parts: p(n,k) = p(n-k,k) + p(n-1,k-1)
 * (integer_partition.cpp:20-26) */
static int64_t exact_parts(int n, int k)
{
    if (n == 0 && k == 0)
        return 1;
    if (n <= 0 || k <= 0)
        return 0;
    return exact_parts(n - k, k) + exact_parts(n - 1, k - 1);
}

/* integer_partition.cpp:30-38 */
int64_t oracle_partition_count(int n, int s)
{
    int64_t total = 0;
    for (int p = 1; p <= s; ++p)
        total += exact_parts(n, p);
    return total;
}

/* Copy the current partition x[1..len] into row r of out (s columns, zero padded). */
static void emit_partition(const int32_t* x, int len, int s, int32_t* out, int64_t* r)
{
    int32_t* row = out + (*r) * s;
    int p = 0;
    for (; p < len; ++p)
        row[p] = x[p + 1];
    for (; p < s; ++p)
        row[p] = 0;
    ++(*r);
}

/* The reference enumerates the partitions of n starting from 1+1+...+1 and moving to
 * the next partition in place (the x[], h, m, r bookkeeping of
 * integer_partition.cpp:41-105); partitions with more than s parts are skipped.
 * The same bookkeeping is restated here. */
int64_t oracle_partitions(int n, int s, int32_t* out)
{
    int64_t rows = 0;
    int32_t* x = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 2));
    for (int i = 1; i <= n; ++i)
        x[i] = 1;
    if (n <= s)
        emit_partition(x, n, s, out, &rows); /* all ones (integer_partition.cpp:58-62) */
    if (n > 1) {
        int h = 1, m = n - 1;
        x[0] = -1;
        x[1] = 2;
        if (m <= s)
            emit_partition(x, m, s, out, &rows);
        while (x[1] != n) {
            if (m - h > 1) { /* split: append another 2 (integer_partition.cpp:76-80) */
                h += 1;
                x[h] = 2;
                m -= 1;
            } else { /* carry (integer_partition.cpp:80-94) */
                int j = m - 2;
                while (x[j] == x[m - 1]) {
                    x[j] = 1;
                    j -= 1;
                }
                h = j + 1;
                x[h] = x[m - 1] + 1;
                int r = x[m] + x[m - 1] * (m - h - 1);
                x[m] = 1;
                if (m - h > 1)
                    x[m - 1] = 1;
                m = h + r - 1;
            }
            if (m <= s)
                emit_partition(x, m, s, out, &rows);
        }
    }
    free(x);
    return rows;
}

/* ======================================================================== */
/* hierarchy graph (src/heom/hierarchy_graph.cpp:30-181)                     */
/* ======================================================================== */

/* binomial coefficient evaluated in double and rounded, as the reference does with
 * boost::math::binomial_coefficient<double> + llround (hierarchy_graph.cpp:34,40) */
static int64_t binom_rounded(int n, int k)
{
    if (k < 0 || k > n)
        return 0;
    if (k > n - k)
        k = n - k;
    long double v = 1.0L;
    for (int i = 1; i <= k; ++i)
        v = v * (long double)(n - k + i) / (long double)i;
    return (int64_t)llroundl(v);
}

int64_t oracle_graph_nodes(int width, int depth) { return binom_rounded(width + depth, width); }

int64_t oracle_graph_first_non_plus(int width, int depth)
{
    return binom_rounded(width + depth - 1, width);
}

/* lexicographic successor of a[0..n) (same contract as std::next_permutation):
 * returns 0 and leaves a sorted ascending when a was the last permutation */
static int next_permutation_i32(int32_t* a, int n)
{
    int i = n - 2;
    while (i >= 0 && a[i] >= a[i + 1])
        --i;
    if (i < 0) {
        for (int l = 0, r = n - 1; l < r; ++l, --r) {
            int32_t t = a[l];
            a[l] = a[r];
            a[r] = t;
        }
        return 0;
    }
    int j = n - 1;
    while (a[j] <= a[i])
        --j;
    int32_t t = a[i];
    a[i] = a[j];
    a[j] = t;
    for (int l = i + 1, r = n - 1; l < r; ++l, --r) {
        t = a[l];
        a[l] = a[r];
        a[r] = t;
    }
    return 1;
}

static int cmp_i32_asc(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* open-addressing map tuple -> uid (the reference uses an unordered_map,
 * hierarchy_graph.cpp:49) */
typedef struct {
    int64_t cap;
    int width;
    int64_t* slot; /* uid or -1 */
    const int32_t* tuples;
} tuple_map;

static uint64_t hash_tuple(const int32_t* t, int w)
{
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < w; ++i) {
        h ^= (uint64_t)(uint32_t)t[i];
        h *= 1099511628211ULL;
        h ^= h >> 29;
    }
    return h;
}

static int64_t map_find(const tuple_map* mp, const int32_t* key)
{
    uint64_t h = hash_tuple(key, mp->width) & (uint64_t)(mp->cap - 1);
    for (;;) {
        int64_t u = mp->slot[h];
        if (u < 0)
            return -1;
        if (memcmp(mp->tuples + u * mp->width, key, sizeof(int32_t) * (size_t)mp->width) == 0)
            return u;
        h = (h + 1) & (uint64_t)(mp->cap - 1);
    }
}

static void map_insert(tuple_map* mp, int64_t uid)
{
    const int32_t* key = mp->tuples + uid * mp->width;
    uint64_t h = hash_tuple(key, mp->width) & (uint64_t)(mp->cap - 1);
    while (mp->slot[h] >= 0)
        h = (h + 1) & (uint64_t)(mp->cap - 1);
    mp->slot[h] = uid;
}

int oracle_graph_build(int width, int depth, int32_t* tuples, int32_t* plus, int32_t* minus,
                       int64_t* nodes_per_depth)
{
    const int64_t nodes = oracle_graph_nodes(width, depth);
    const int32_t invalid = -2; /* hierarchy_graph.cpp:24 */
    int64_t uid = 0;

    /* Step 1: nodes layer by layer; inside a layer, partitions in generator order,
     * each sorted ascending and expanded into all of its permutations
     * (hierarchy_graph.cpp:57-117) */
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)width);
    for (int d = 0; d <= depth; ++d) {
        int64_t count_d = 0;
        if (d == 0) {
            memset(tuples, 0, sizeof(int32_t) * (size_t)width);
            uid = 1;
            count_d = 1;
        } else {
            int64_t np = oracle_partition_count(d, width);
            int32_t* parts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np * width));
            int64_t got = oracle_partitions(d, width, parts);
            if (got != np) {
                free(parts);
                free(cur);
                return -1;
            }
            for (int64_t p = 0; p < np; ++p) {
                memcpy(cur, parts + p * width, sizeof(int32_t) * (size_t)width);
                qsort(cur, (size_t)width, sizeof(int32_t), cmp_i32_asc);
                do {
                    if (uid >= nodes) {
                        free(parts);
                        free(cur);
                        return -2;
                    }
                    memcpy(tuples + uid * width, cur, sizeof(int32_t) * (size_t)width);
                    ++uid;
                    ++count_d;
                } while (next_permutation_i32(cur, width));
            }
            free(parts);
        }
        if (nodes_per_depth)
            nodes_per_depth[d] = count_d;
    }
    free(cur);
    if (uid != nodes)
        return -3;

    /* Step 2: plus / minus edges by tuple lookup (hierarchy_graph.cpp:122-168) */
    tuple_map mp;
    mp.width = width;
    mp.tuples = tuples;
    mp.cap = 1;
    while (mp.cap < 2 * nodes + 16)
        mp.cap <<= 1;
    mp.slot = (int64_t*)malloc(sizeof(int64_t) * (size_t)mp.cap);
    for (int64_t i = 0; i < mp.cap; ++i)
        mp.slot[i] = -1;
    for (int64_t u = 0; u < nodes; ++u)
        map_insert(&mp, u);

    int err = 0;
#pragma omp parallel if (nodes > 4096)
    {
        int32_t* probe = (int32_t*)malloc(sizeof(int32_t) * (size_t)width);
#pragma omp for schedule(static)
        for (int64_t u = 0; u < nodes; ++u) {
            const int32_t* t = tuples + u * width;
            int dep = 0;
            for (int m = 0; m < width; ++m)
                dep += t[m];
            for (int m = 0; m < width; ++m) {
                memcpy(probe, t, sizeof(int32_t) * (size_t)width);
                int32_t pu = invalid, mu = invalid;
                if (dep < depth) {
                    probe[m] += 1;
                    int64_t f = map_find(&mp, probe);
                    if (f < 0)
                        err = 1;
                    pu = (int32_t)f;
                    probe[m] -= 1;
                }
                if (t[m] - 1 >= 0) {
                    probe[m] -= 1;
                    int64_t f = map_find(&mp, probe);
                    if (f < 0)
                        err = 1;
                    mu = (int32_t)f;
                }
                plus[u * width + m] = pu;
                minus[u * width + m] = mu;
            }
        }
        free(probe);
    }
    free(mp.slot);
    return err ? -4 : 0;
}

/* ======================================================================== */
/* precompute (src/heom/instance.cpp:155-262)                                */
/* ======================================================================== */

void oracle_precompute(int baths, int matsubaras, const double* lambda, const double* invnu,
                       const double* omega, double temperature, int64_t matrices, int width,
                       const int32_t* tuples, double* pf_same_sum, double* pf_same,
                       double* pf_circ, double* cbk_out, double* cbk_sqrt)
{
    const double hbar = ORACLE_HBAR;
    const double beta = 1.0 / (ORACLE_KB * temperature); /* instance.cpp:167 */
    const int row = (matsubaras - 1) * ORACLE_MAX_PADE;

    for (int b = 0; b < baths; ++b) {
        double complex lambda_b = CMPLX(lambda[b], 0.0);
        double complex nu_b = CMPLX(1.0 / invnu[b], omega[b]); /* instance.cpp:174 */
        double complex circ = nu_b * lambda_b / hbar;          /* instance.cpp:176 */
        double complex aux = 2.0 * lambda_b / (beta * hbar * hbar * nu_b); /* :177 */

        for (int k = 0; k < matsubaras; ++k) {
            double complex c, nubk;
            if (k == 0) { /* Drude term with the Pade correction (instance.cpp:183-200) */
                double complex auxsum = 0.0;
                for (int j = 0; j < matsubaras; ++j) {
                    double eta = (double)oracle_eta_pade[row + j];
                    double xi = (double)oracle_xi_pade[row + j];
                    auxsum += (2.0 * eta * nu_b * nu_b) /
                              (xi * xi / (beta * beta * hbar * hbar) - nu_b * nu_b);
                }
                double complex c0 = 2.0 * lambda_b / (beta * hbar) * (1.0 - auxsum);
                c = c0;
                nubk = nu_b;
            } else { /* Pade pole k (instance.cpp:202-214) */
                double eta = (double)oracle_eta_pade[row + k];
                double xi = (double)oracle_xi_pade[row + k];
                c = 4.0 * lambda_b * nu_b / (beta * beta * hbar * hbar) * (eta * xi) /
                    (xi * xi / (beta * beta * hbar * hbar) - nu_b * nu_b);
                nubk = xi / (beta * hbar);
            }
            cbk_out[2 * (b * matsubaras + k)] = creal(c);      /* instance.cpp:216 */
            cbk_out[2 * (b * matsubaras + k) + 1] = cimag(c);
            cbk_sqrt[b * matsubaras + k] = sqrt(cabs(c));      /* instance.cpp:217 */
            aux -= c / (hbar * nubk);                           /* instance.cpp:218 */
        }
        pf_same[2 * b] = creal(aux);
        pf_same[2 * b + 1] = cimag(aux);
        pf_circ[2 * b] = creal(circ);
        pf_circ[2 * b + 1] = cimag(circ);
    }

    /* S_n = -sum_{b,k} n_bk nu_bk in long double (instance.cpp:227-262) */
#pragma omp parallel for schedule(static) if (matrices > 1024)
    for (int64_t id = 0; id < matrices; ++id) {
        long double zr = 0.0L, zi = 0.0L;
        for (int b = 0; b < baths; ++b) {
            double complex nu_b = CMPLX(1.0 / invnu[b], omega[b]);
            for (int k = 0; k < matsubaras; ++k) {
                long double nr, ni;
                if (k == 0) {
                    nr = (long double)creal(nu_b);
                    ni = (long double)cimag(nu_b);
                } else {
                    nr = oracle_xi_pade[row + k] / (long double)(beta * hbar);
                    ni = 0.0L;
                }
                int m = b * matsubaras + k;
                zr -= nr * (long double)tuples[id * width + m];
                zi -= ni * (long double)tuples[id * width + m];
            }
        }
        pf_same_sum[2 * id] = (double)zr;
        pf_same_sum[2 * id + 1] = (double)zi;
    }
}

/* ======================================================================== */
/* RHS (cl/heom/heom_ode.cl:24-451)                                          */
/* ======================================================================== */

/* tss = 1: heom_ode.cl compiled with HEOM_TSS_EQUATION (:116-121, :130-136) */
static void rhs_eval(const oracle_instance* inst, const double* in, double* out, int tss)
{
    const int N = inst->states;
    const int K = inst->matsubaras;
    const int BM = inst->baths_max_per_state;
    const int W = inst->width;
    const int64_t M = inst->matrices;
    const double HB = ORACLE_HBAR;
    const double* H = inst->hamiltonian;

#pragma omp parallel for schedule(dynamic, 64) if (M * N * N > 16384)
    for (int64_t id = 0; id < M; ++id) {
        const double* s = in + id * N * N * 2;
        double* o = out + id * N * N * 2;
        for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) {
                const double s_r = s[(i * N + j) * 2];
                const double s_i = s[(i * N + j) * 2 + 1];

                /* commutator (i/ħ)(σH - Hσ) with separate temporaries (:78-114); for the
                 * thermal state search the elementwise product σ_ij H_ij (:116-121) */
                double com_r = 0.0, com_i = 0.0;
                if (tss) {
                    com_r = s_r * H[(i * N + j) * 2] - s_i * H[(i * N + j) * 2 + 1];
                    com_i = s_r * H[(i * N + j) * 2 + 1] + s_i * H[(i * N + j) * 2];
                }
                for (int k = 0; k < N && !tss; ++k) {
                    double sh_r = s[(i * N + k) * 2] * H[(k * N + j) * 2];
                    sh_r -= s[(i * N + k) * 2 + 1] * H[(k * N + j) * 2 + 1];
                    double sh_i = s[(i * N + k) * 2] * H[(k * N + j) * 2 + 1];
                    sh_i += s[(i * N + k) * 2 + 1] * H[(k * N + j) * 2];
                    double hs_r = H[(i * N + k) * 2] * s[(k * N + j) * 2];
                    hs_r -= H[(i * N + k) * 2 + 1] * s[(k * N + j) * 2 + 1];
                    double hs_i = H[(i * N + k) * 2] * s[(k * N + j) * 2 + 1];
                    hs_i += H[(i * N + k) * 2 + 1] * s[(k * N + j) * 2];
                    com_r += sh_r - hs_r;
                    com_i += sh_i - hs_i;
                }
                if (!tss) {
                    double t = com_r;
                    com_r = -com_i / HB;
                    com_i = t / HB;
                }

                /* same term: S_n σ_ij (not for the thermal state search) and -Δ_b for the
                 * baths of i and of j (:126-169) */
                const double pss_r = tss ? 0.0 : inst->pf_same_sum[2 * id];
                const double pss_i = tss ? 0.0 : inst->pf_same_sum[2 * id + 1];
                double same_r = pss_r * s_r - pss_i * s_i;
                double same_i = pss_r * s_i + pss_i * s_r;
                if (i != j) {
                    for (int b = 0; b < BM; ++b) {
                        const int bi = inst->coupling[i * BM + b];
                        if (bi >= 0) {
                            const double mr = -inst->pf_same[2 * bi];
                            const double mi = -inst->pf_same[2 * bi + 1];
                            same_r += mr * s_r - mi * s_i;
                            same_i += mr * s_i + mi * s_r;
                        }
                        const int bj = inst->coupling[j * BM + b];
                        if (bj >= 0) {
                            const double mr = -inst->pf_same[2 * bj];
                            const double mi = -inst->pf_same[2 * bj + 1];
                            same_r += mr * s_r - mi * s_i;
                            same_i += mr * s_i + mi * s_r;
                        }
                    }
                }

                /* plus term, separate i and j accumulators (:172-266) */
                double pl_ir = 0.0, pl_ii = 0.0, pl_jr = 0.0, pl_ji = 0.0;
                if (id < inst->first_non_plus) {
                    for (int b = 0; b < BM; ++b) {
                        const int bi = inst->coupling[i * BM + b];
                        if (bi >= 0) {
                            for (int m = 0; m < K; ++m) {
                                const int mm = bi * K + m;
                                const int32_t pu = inst->plus[id * W + mm];
                                if (!inst->mask || inst->mask[pu]) {
                                    const double* sp = in + (int64_t)pu * N * N * 2;
                                    const double sp_r = sp[(i * N + j) * 2];
                                    const double sp_i = sp[(i * N + j) * 2 + 1];
                                    const int t = inst->tuples[id * W + mm];
                                    const double a_r = 0.0;
                                    const double a_i =
                                        sqrt((t + 1.0) * hypot(inst->cbk[2 * mm], inst->cbk[2 * mm + 1]) / HB);
                                    pl_ir += a_r * sp_r - a_i * sp_i;
                                    pl_ii += a_r * sp_i + a_i * sp_r;
                                }
                            }
                        }
                        const int bj = inst->coupling[j * BM + b];
                        if (bj >= 0) {
                            for (int m = 0; m < K; ++m) {
                                const int mm = bj * K + m;
                                const int32_t pu = inst->plus[id * W + mm];
                                if (!inst->mask || inst->mask[pu]) {
                                    const double* sp = in + (int64_t)pu * N * N * 2;
                                    const double sp_r = sp[(i * N + j) * 2];
                                    const double sp_i = sp[(i * N + j) * 2 + 1];
                                    const int t = inst->tuples[id * W + mm];
                                    const double a_r = 0.0;
                                    const double a_i =
                                        sqrt((t + 1.0) * hypot(inst->cbk[2 * mm], inst->cbk[2 * mm + 1]) / HB);
                                    pl_jr -= a_r * sp_r - a_i * sp_i;
                                    pl_ji -= a_r * sp_i + a_i * sp_r;
                                }
                            }
                        }
                    }
                }
                const double plus_r = pl_ir + pl_jr;
                const double plus_i = pl_ii + pl_ji;

                /* minus term (:269-409) */
                double mn_ir = 0.0, mn_ii = 0.0, mn_jr = 0.0, mn_ji = 0.0;
                for (int b = 0; b < BM; ++b) {
                    const int bi = inst->coupling[i * BM + b];
                    if (bi >= 0) {
                        for (int m = 0; m < K; ++m) {
                            const int mm = bi * K + m;
                            const int t = inst->tuples[id * W + mm];
                            if (t > 0) {
                                const int32_t mu = inst->minus[id * W + mm];
                                if (!inst->mask || inst->mask[mu]) {
                                    const double* sm = in + (int64_t)mu * N * N * 2;
                                    const double sm_r = sm[(i * N + j) * 2];
                                    const double sm_i = sm[(i * N + j) * 2 + 1];
                                    const double c_r = inst->cbk[2 * mm];
                                    const double c_i = inst->cbk[2 * mm + 1];
                                    const double pc_r = sm_r * c_r - sm_i * c_i;
                                    const double pc_i = sm_r * c_i + sm_i * c_r;
                                    const double a_r = -pc_i / HB;
                                    const double a_i = pc_r / HB;
                                    const double sc = sqrt(t / hypot(c_r, c_i) * HB);
                                    mn_ir += a_r * sc;
                                    mn_ii += a_i * sc;
                                    if (m == 0) {
                                        const double f_r = inst->pf_circ[2 * bi];
                                        const double f_i = inst->pf_circ[2 * bi + 1];
                                        mn_ir += (sm_r * f_r - sm_i * f_i) * sc;
                                        mn_ii += (sm_r * f_i + sm_i * f_r) * sc;
                                    }
                                }
                            }
                        }
                    }
                    const int bj = inst->coupling[j * BM + b];
                    if (bj >= 0) {
                        for (int m = 0; m < K; ++m) {
                            const int mm = bj * K + m;
                            const int t = inst->tuples[id * W + mm];
                            if (t > 0) {
                                const int32_t mu = inst->minus[id * W + mm];
                                if (!inst->mask || inst->mask[mu]) {
                                    const double* sm = in + (int64_t)mu * N * N * 2;
                                    const double sm_r = sm[(i * N + j) * 2];
                                    const double sm_i = sm[(i * N + j) * 2 + 1];
                                    const double c_r = inst->cbk[2 * mm];
                                    const double c_i = inst->cbk[2 * mm + 1];
                                    const double pc_r = sm_r * c_r - sm_i * c_i;
                                    const double pc_i = sm_r * c_i + sm_i * c_r;
                                    const double a_r = pc_i / HB;
                                    const double a_i = -pc_r / HB;
                                    const double sc = sqrt(t / hypot(c_r, c_i) * HB);
                                    mn_jr += a_r * sc;
                                    mn_ji += a_i * sc;
                                    if (m == 0) {
                                        const double f_r = inst->pf_circ[2 * bj];
                                        const double f_i = inst->pf_circ[2 * bj + 1];
                                        mn_jr += (sm_r * f_r - sm_i * f_i) * sc;
                                        mn_ji += (sm_r * f_i + sm_i * f_r) * sc;
                                    }
                                }
                            }
                        }
                    }
                }
                const double minus_r = mn_ir + mn_jr;
                const double minus_i = mn_ii + mn_ji;

                /* sum of the four groups (:413-414), pure evaluation (:432-433) */
                o[(i * N + j) * 2] = com_r + same_r + plus_r + minus_r;
                o[(i * N + j) * 2 + 1] = com_i + same_i + plus_i + minus_i;
            }
        }
    }
}

void oracle_rhs(const oracle_instance* inst, const double* in, double* out) { rhs_eval(inst, in, out, 0); }

void oracle_rhs_tss(const oracle_instance* inst, const double* in, double* out) { rhs_eval(inst, in, out, 1); }

/* ======================================================================== */
/* RK (cl/heom/rk_weighted_add.cl:19-106 and the classical RK4 tableau)      */
/* ======================================================================== */

void oracle_rk_weighted_add(int64_t size, int n, double h, double* out, const double* y_n,
                            const double* c, const double* const* k)
{
#pragma omp parallel for schedule(static) if (size > 65536)
    for (int64_t e = 0; e < size; ++e) {
        double v = 0.0;
        for (int l = 0; l < n; ++l)
            if (c[l] != 0.0)
                v += c[l] * k[l][e];
        v *= h;
        v += y_n[e];
        out[e] = v;
    }
}

/* Classical RK4 (Butcher a = [[1/2], [0, 1/2], [0, 0, 1]], b = [1/6, 1/3, 1/3, 1/6]).
 * The reference drives it through the un-vendored noma/num stepper
 * (include/heom/population_dynamics_solver.hpp:112-114, 242); here every stage input
 * and the final combination go through rk_weighted_add as written above. */
void oracle_rk4_step(const oracle_instance* inst, double h, double* y, double* work)
{
    const int64_t size = inst->matrices * inst->states * inst->states * 2;
    double* k1 = work;
    double* k2 = work + size;
    double* k3 = work + 2 * size;
    double* k4 = work + 3 * size;
    double* t = work + 4 * size;
    const double* ks[4] = {k1, k2, k3, k4};
    const double a2[1] = {0.5};
    const double a3[2] = {0.0, 0.5};
    const double a4[3] = {0.0, 0.0, 1.0};
    const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

    oracle_rhs(inst, y, k1);
    oracle_rk_weighted_add(size, 1, h, t, y, a2, ks);
    oracle_rhs(inst, t, k2);
    oracle_rk_weighted_add(size, 2, h, t, y, a3, ks);
    oracle_rhs(inst, t, k3);
    oracle_rk_weighted_add(size, 3, h, t, y, a4, ks);
    oracle_rhs(inst, t, k4);
    oracle_rk_weighted_add(size, 4, h, t, y, b, ks);
    memcpy(y, t, sizeof(double) * (size_t)size);
}

int64_t oracle_propagate(const oracle_instance* inst, double h, int64_t nsteps,
                         int64_t observe_steps, double* y, double* work, double* tops)
{
    const int64_t top = (int64_t)inst->states * inst->states * 2;
    int64_t rows = 0;
    if (tops)
        memcpy(tops + (rows++) * top, y, sizeof(double) * (size_t)top);
    for (int64_t s = 1; s <= nsteps; ++s) {
        oracle_rk4_step(inst, h, y, work);
        if (tops && observe_steps > 0 && s % observe_steps == 0)
            memcpy(tops + (rows++) * top, y, sizeof(double) * (size_t)top);
    }
    return rows;
}

/* ======================================================================== */
/* observers and hierarchy arithmetic                                        */
/* ======================================================================== */

void oracle_hierarchy_norm(int states, int64_t matrices, const double* sigma, double threshold,
                           double* norms, uint8_t* mask)
{
#pragma omp parallel for schedule(static) if (matrices * states * states > 16384)
    for (int64_t id = 0; id < matrices; ++id) {
        const double* a = sigma + id * states * states * 2;
        double acc = 0.0;
        for (int e = 0; e < states * states; ++e)
            acc += a[2 * e] * a[2 * e] + a[2 * e + 1] * a[2 * e + 1];
        norms[id] = acc;
        if (mask && acc - threshold <= 0.0 && mask[id] != 2)
            mask[id] = 0;
    }
}

void oracle_mmult_left(int states, int64_t matrices, const double* mx, double* sigma)
{
    const int N = states;
#pragma omp parallel if (matrices * N * N > 16384)
    {
        double* tmp = (double*)malloc(sizeof(double) * (size_t)(N * N * 2));
#pragma omp for schedule(static)
        for (int64_t id = 0; id < matrices; ++id) {
            double* a = sigma + id * N * N * 2;
            for (int i = 0; i < N; ++i) {
                for (int j = 0; j < N; ++j) {
                    double vr = 0.0, vi = 0.0;
                    for (int k = 0; k < N; ++k) {
                        const double mr = mx[(i * N + k) * 2], mi = mx[(i * N + k) * 2 + 1];
                        const double ar = a[(k * N + j) * 2], ai = a[(k * N + j) * 2 + 1];
                        const double pr = mr * ar - mi * ai;
                        const double pi = mr * ai + mi * ar;
                        vr = vr + pr;
                        vi = vi + pi;
                    }
                    tmp[(i * N + j) * 2] = vr;
                    tmp[(i * N + j) * 2 + 1] = vi;
                }
            }
            memcpy(a, tmp, sizeof(double) * (size_t)(N * N * 2));
        }
        free(tmp);
    }
}

void oracle_trace(int states, const double* d, const double* top, double* trace2)
{
    const int N = states;
    double tr = 0.0, ti = 0.0;
    for (int i = 0; i < N; ++i) {
        double vr, vi;
        if (!d) {
            vr = top[(i * N + i) * 2];
            vi = top[(i * N + i) * 2 + 1];
        } else {
            vr = 0.0;
            vi = 0.0;
            for (int k = 0; k < N; ++k) {
                const double dr = d[(i * N + k) * 2], di = d[(i * N + k) * 2 + 1];
                const double rr = top[(k * N + i) * 2], ri = top[(k * N + i) * 2 + 1];
                vr += dr * rr - di * ri;
                vi += dr * ri + di * rr;
            }
        }
        tr += vr;
        ti += vi;
    }
    trace2[0] = tr;
    trace2[1] = ti;
}

/* ======================================================================== */
/* adjoint: transpose of the RHS and gradients of observables               */
/* ======================================================================== */
/* The reference has no differentiation; these functions restate the mathematics of
 * the discrete adjoint of oracle_rk4_step with oracle_rhs as the operator L.  The RHS is
 * complex-linear in the state, so the transpose of L as a map of the interleaved real
 * vector is the complex adjoint L^dag under <a, b> = Re sum conj(a) b. */

static inline const double* in_state(const double* y, int64_t id, int N) { return y + id * N * N * 2; }

/* out = L^dag mu.  Pass 1 transposes the per-ADO parts of oracle_rhs (commutator and
 * same term); pass 2 scatters every neighbour term of oracle_rhs back onto the ADO it
 * read, one element plane per thread, with the conjugate coefficient. */
void oracle_rhs_adjoint(const oracle_instance* inst, const double* mu, double* out)
{
    const int N = inst->states;
    const int K = inst->matsubaras;
    const int BM = inst->baths_max_per_state;
    const int W = inst->width;
    const int64_t M = inst->matrices;
    const double HB = ORACLE_HBAR;
    const double* H = inst->hamiltonian;

#pragma omp parallel for schedule(dynamic, 64) if (M * N * N > 16384)
    for (int64_t id = 0; id < M; ++id) {
        const double* m = mu + id * N * N * 2;
        double* o = out + id * N * N * 2;
        for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) {
                /* the forward map s -> (i/hbar)(s H - H s) has the adjoint
                 * mu -> (-i/hbar)(mu H^dag - H^dag mu) */
                double ar = 0.0, ai = 0.0;
                for (int k = 0; k < N; ++k) {
                    const double mr = m[(i * N + k) * 2], mi = m[(i * N + k) * 2 + 1];
                    const double hr = H[(j * N + k) * 2], hi = -H[(j * N + k) * 2 + 1]; /* conj(H_jk) */
                    ar += mr * hr - mi * hi;
                    ai += mr * hi + mi * hr;
                    const double gr = H[(k * N + i) * 2], gi = -H[(k * N + i) * 2 + 1]; /* conj(H_ki) */
                    const double nr = m[(k * N + j) * 2], ni = m[(k * N + j) * 2 + 1];
                    ar -= gr * nr - gi * ni;
                    ai -= gr * ni + gi * nr;
                }
                double o_r = ai / HB, o_i = -ar / HB;

                /* same term: multiplication by conj(S_n - sum Delta) */
                double sr = inst->pf_same_sum[2 * id], si = inst->pf_same_sum[2 * id + 1];
                if (i != j) {
                    for (int b = 0; b < BM; ++b) {
                        const int bi = inst->coupling[i * BM + b];
                        if (bi >= 0) {
                            sr -= inst->pf_same[2 * bi];
                            si -= inst->pf_same[2 * bi + 1];
                        }
                        const int bj = inst->coupling[j * BM + b];
                        if (bj >= 0) {
                            sr -= inst->pf_same[2 * bj];
                            si -= inst->pf_same[2 * bj + 1];
                        }
                    }
                }
                const double mr = m[(i * N + j) * 2], mi = m[(i * N + j) * 2 + 1];
                o_r += sr * mr + si * mi;
                o_i += sr * mi - si * mr;
                o[(i * N + j) * 2] = o_r;
                o[(i * N + j) * 2 + 1] = o_i;
            }
        }
    }

#pragma omp parallel for schedule(static) if (N * N > 64)
    for (int e = 0; e < N * N; ++e) {
        const int i = e / N, j = e % N;
        for (int64_t id = 0; id < M; ++id) {
            const double mr = mu[(id * N * N + e) * 2], mi = mu[(id * N * N + e) * 2 + 1];
            for (int side = 0; side < 2; ++side) {
                const int s = side == 0 ? i : j;
                const double sg = side == 0 ? 1.0 : -1.0; /* +i for state i, -i for state j */
                for (int b = 0; b < BM; ++b) {
                    const int bs = inst->coupling[s * BM + b];
                    if (bs < 0)
                        continue;
                    for (int m = 0; m < K; ++m) {
                        const int mm = bs * K + m;
                        const int t = inst->tuples[id * W + mm];
                        const double cr = inst->cbk[2 * mm], ci = inst->cbk[2 * mm + 1];
                        /* plus edge: forward sg * i * a * sigma[pu] added to ADO id */
                        if (id < inst->first_non_plus) {
                            const int32_t pu = inst->plus[id * W + mm];
                            if (!inst->mask || inst->mask[pu]) {
                                const double a = sqrt((t + 1.0) * hypot(cr, ci) / HB);
                                /* conj(sg i a) mu = -sg i a mu */
                                double* op = out + ((int64_t)pu * N * N + e) * 2;
                                op[0] += sg * a * mi;
                                op[1] -= sg * a * mr;
                            }
                        }
                        /* minus edge: forward z * sigma[mu] with
                         * z = (sg i c / hbar + [m == 0] Phi) sqrt(t hbar / |c|) */
                        if (t > 0) {
                            const int32_t mu_id = inst->minus[id * W + mm];
                            if (!inst->mask || inst->mask[mu_id]) {
                                const double sc = sqrt(t / hypot(cr, ci) * HB);
                                double zr = -sg * ci / HB, zi = sg * cr / HB;
                                if (m == 0) {
                                    zr += inst->pf_circ[2 * bs];
                                    zi += inst->pf_circ[2 * bs + 1];
                                }
                                zr *= sc;
                                zi *= sc;
                                double* op = out + ((int64_t)mu_id * N * N + e) * 2;
                                op[0] += zr * mr + zi * mi;
                                op[1] += zr * mi - zi * mr;
                            }
                        }
                    }
                }
            }
        }
    }
}

/* g += w * mu conj(B) for a coefficient kappa that enters the RHS as kappa * B */
static inline void grad_linear(double* g, double w, double mr, double mi, double br, double bi)
{
    g[0] += w * (mr * br + mi * bi);
    g[1] += w * (mi * br - mr * bi);
}

/* Adds w * d<mu, L y>/d(kappa) for every coefficient array kappa of the instance, each
 * complex entry as dRe + i dIm: g_h [2 N N], g_pss [2M], g_same [2B], g_circ [2B],
 * g_cbk [2BK].  Every term of oracle_rhs is differentiated where it is written: the
 * commutator, S_n and Delta_b enter linearly, Phi_b linearly in the minus term, and c
 * through c itself and |c| in the plus and minus amplitudes.  Modes with c = 0 are
 * skipped (their amplitudes are not differentiable). */
void oracle_adjoint_contract(const oracle_instance* inst, const double* mu, const double* y, double w,
                             double* g_h, double* g_pss, double* g_same, double* g_circ, double* g_cbk)
{
    const int N = inst->states;
    const int K = inst->matsubaras;
    const int BM = inst->baths_max_per_state;
    const int W = inst->width;
    const int B = W / K;
    const int64_t M = inst->matrices;
    const double HB = ORACLE_HBAR;

#pragma omp parallel if (M * N * N > 16384)
    {
        double* lh = calloc((size_t)(2 * N * N + 4 * B + 2 * W), sizeof(double));
        double* ls = lh + 2 * N * N;
        double* lc = ls + 2 * B;
        double* lk = lc + 2 * B;
#pragma omp for schedule(dynamic, 64)
        for (int64_t id = 0; id < M; ++id) {
            const double* m = mu + id * N * N * 2;
            const double* s = y + id * N * N * 2;
            double gp_r = 0.0, gp_i = 0.0;
            for (int i = 0; i < N; ++i) {
                for (int j = 0; j < N; ++j) {
                    const double mr = m[(i * N + j) * 2], mi = m[(i * N + j) * 2 + 1];
                    const double yr = s[(i * N + j) * 2], yi = s[(i * N + j) * 2 + 1];
                    /* commutator (i/hbar)(sum_k y_ik H_kj - H_ik y_kj) */
                    for (int k = 0; k < N; ++k) {
                        const double ar = s[(i * N + k) * 2], ai = s[(i * N + k) * 2 + 1];
                        grad_linear(lh + 2 * (k * N + j), w, mr, mi, -ai / HB, ar / HB);
                        const double br = s[(k * N + j) * 2], bi = s[(k * N + j) * 2 + 1];
                        grad_linear(lh + 2 * (i * N + k), w, mr, mi, bi / HB, -br / HB);
                    }
                    /* S_n y_ij */
                    gp_r += w * (mr * yr + mi * yi);
                    gp_i += w * (mi * yr - mr * yi);
                    /* -Delta_b y_ij for the baths of i and j, i != j */
                    if (i != j) {
                        for (int b = 0; b < BM; ++b) {
                            const int bi = inst->coupling[i * BM + b];
                            if (bi >= 0)
                                grad_linear(ls + 2 * bi, w, mr, mi, -yr, -yi);
                            const int bj = inst->coupling[j * BM + b];
                            if (bj >= 0)
                                grad_linear(ls + 2 * bj, w, mr, mi, -yr, -yi);
                        }
                    }
                    for (int side = 0; side < 2; ++side) {
                        const int st = side == 0 ? i : j;
                        const double sg = side == 0 ? 1.0 : -1.0;
                        for (int b = 0; b < BM; ++b) {
                            const int bs = inst->coupling[st * BM + b];
                            if (bs < 0)
                                continue;
                            for (int k = 0; k < K; ++k) {
                                const int mm = bs * K + k;
                                const int t = inst->tuples[id * W + mm];
                                const double cr = inst->cbk[2 * mm], ci = inst->cbk[2 * mm + 1];
                                const double ac = hypot(cr, ci);
                                if (ac == 0.0)
                                    continue;
                                if (id < inst->first_non_plus) {
                                    const int32_t pu = inst->plus[id * W + mm];
                                    if (!inst->mask || inst->mask[pu]) {
                                        const double* p = in_state(y, pu, N);
                                        const double pr = p[(i * N + j) * 2], pi = p[(i * N + j) * 2 + 1];
                                        /* term sg i a p, a = sqrt((t+1)|c|/hbar), da/dx = a x / (2 |c|^2) */
                                        const double a = sqrt((t + 1.0) * ac / HB);
                                        const double da = a / (2.0 * ac * ac);
                                        /* d term / dx = sg i da x p: Re(conj(mu) i p) = mr (-pi) ... */
                                        const double q = -(mr * pi) + mi * pr; /* Re(conj(mu) * i p) */
                                        lk[2 * mm] += w * sg * da * cr * q;
                                        lk[2 * mm + 1] += w * sg * da * ci * q;
                                    }
                                }
                                if (t > 0) {
                                    const int32_t mu_id = inst->minus[id * W + mm];
                                    if (!inst->mask || inst->mask[mu_id]) {
                                        const double* p = in_state(y, mu_id, N);
                                        const double pr = p[(i * N + j) * 2], pi = p[(i * N + j) * 2 + 1];
                                        const double sc = sqrt(t / ac * HB);
                                        const double dsc = -sc / (2.0 * ac * ac); /* dsc/dx = dsc * x */
                                        /* z = sg i c / hbar + [k == 0] Phi_b */
                                        double zr = -sg * ci / HB, zi = sg * cr / HB;
                                        if (k == 0) {
                                            zr += inst->pf_circ[2 * bs];
                                            zi += inst->pf_circ[2 * bs + 1];
                                            /* Phi_b enters as Phi_b * sc * p */
                                            grad_linear(lc + 2 * bs, w, mr, mi, sc * pr, sc * pi);
                                        }
                                        /* Re(conj(mu) v p) for v = x + i y */
                                        const double re_p = mr * pr + mi * pi;  /* Re(conj(mu) p) */
                                        const double re_ip = mi * pr - mr * pi; /* Re(conj(mu) i p) */
                                        /* d/dcr: sg (i/hbar) sc p + z dsc cr p */
                                        const double d_r = sg * sc / HB * re_ip + dsc * cr * (zr * re_p + zi * re_ip);
                                        /* d/dci: -sg/hbar sc p + z dsc ci p */
                                        const double d_i = -sg * sc / HB * re_p + dsc * ci * (zr * re_p + zi * re_ip);
                                        lk[2 * mm] += w * d_r;
                                        lk[2 * mm + 1] += w * d_i;
                                    }
                                }
                            }
                        }
                    }
                }
            }
            g_pss[2 * id] += gp_r;
            g_pss[2 * id + 1] += gp_i;
        }
#pragma omp critical
        {
            for (int e = 0; e < 2 * N * N; ++e)
                g_h[e] += lh[e];
            for (int e = 0; e < 2 * B; ++e) {
                g_same[e] += ls[e];
                g_circ[e] += lc[e];
            }
            for (int e = 0; e < 2 * W; ++e)
                g_cbk[e] += lk[e];
        }
        free(lh);
    }
}

static void axpy_into(int64_t size, double* out, const double* x, double a, const double* y)
{
#pragma omp parallel for schedule(static) if (size > 65536)
    for (int64_t e = 0; e < size; ++e)
        out[e] = x[e] + a * y[e];
}

/* Tr(O sigma_0) = sum_ij O_ji sigma_0,ij */
static void observe_op(int N, const double* op, const double* y, double* x)
{
    double xr = 0.0, xi = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const double orr = op[(j * N + i) * 2], oi = op[(j * N + i) * 2 + 1];
            const double sr = y[(i * N + j) * 2], si = y[(i * N + j) * 2 + 1];
            xr += orr * sr - oi * si;
            xi += orr * si + oi * sr;
        }
    x[0] = xr;
    x[1] = xi;
}

/* J = sum_r Re(w_r X_r), X_r = Tr(O sigma_0(t_r)), t_r = r * observe_steps * h for
 * r = 0 .. nsteps / observe_steps, over oracle_rk4_step from y0.  Writes X_r into
 * observed and dJ/d(kappa) (dRe + i dIm) for the coefficient arrays as in
 * oracle_adjoint_contract, and for the initial state (g_state, may be NULL).
 *
 * Discrete adjoint of one RK4 step y' = y + h/6 (k1 + 2 k2 + 2 k3 + k4), k_s = L Y_s:
 * with lambda = dJ/dy', the adjoint stages A1 = lambda, A2 = lambda + h/2 L^dag A1,
 * A3 = lambda + h/2 L^dag A2, A4 = lambda + h L^dag A3 give
 * dJ/dy = lambda + h/6 L^dag (A1 + 2 A2 + 2 A3 + A4) and the coefficient gradient
 * h/6 <A4, dL Y1> + h/3 <A3, dL Y2> + h/3 <A2, dL Y3> + h/6 <A1, dL Y4>.
 * The forward states are kept every checkpoint_steps steps and each segment is
 * recomputed before it is swept backwards.  Returns 0, or -1 on bad arguments or
 * allocation failure. */
int oracle_adjoint_gradient(const oracle_instance* inst, double h, int64_t nsteps, int64_t observe_steps,
                            int64_t checkpoint_steps, const double* op, const double* weights, const double* y0,
                            double* observed, double* g_h, double* g_pss, double* g_same, double* g_circ,
                            double* g_cbk, double* g_state)
{
    if (nsteps < 0 || observe_steps < 1 || checkpoint_steps < 1)
        return -1;
    const int N = inst->states;
    const int W = inst->width;
    const int B = W / inst->matsubaras;
    const int64_t M = inst->matrices;
    const int64_t size = M * N * N * 2;
    const int64_t rows = nsteps / observe_steps + 1;
    const int64_t last = (rows - 1) * observe_steps;
    const int64_t ck = checkpoint_steps;
    const int64_t ncp = last / ck + 1;

    double* cps = malloc(sizeof(double) * (size_t)(size * ncp));
    double* seg = malloc(sizeof(double) * (size_t)(size * ck));
    double* wk = malloc(sizeof(double) * (size_t)(size * 16));
    if (!cps || !seg || !wk) {
        free(cps);
        free(seg);
        free(wk);
        return -1;
    }
    double* work = wk;            /* 5 states for oracle_rk4_step */
    double* y = wk + 5 * size;
    double* k = wk + 6 * size;
    double* y2 = wk + 7 * size;
    double* y3 = wk + 8 * size;
    double* y4 = wk + 9 * size;
    double* lam = wk + 10 * size;
    double* a2 = wk + 11 * size;
    double* a3 = wk + 12 * size;
    double* a4 = wk + 13 * size;
    double* f = wk + 14 * size;
    double* acc = wk + 15 * size;

    memset(g_h, 0, sizeof(double) * (size_t)(2 * N * N));
    memset(g_pss, 0, sizeof(double) * (size_t)(2 * M));
    memset(g_same, 0, sizeof(double) * (size_t)(2 * B));
    memset(g_circ, 0, sizeof(double) * (size_t)(2 * B));
    memset(g_cbk, 0, sizeof(double) * (size_t)(2 * W));

    /* forward sweep */
    memcpy(y, y0, sizeof(double) * (size_t)size);
    memcpy(cps, y, sizeof(double) * (size_t)size);
    observe_op(N, op, y, observed);
    for (int64_t s = 1; s <= last; ++s) {
        oracle_rk4_step(inst, h, y, work);
        if (s % observe_steps == 0)
            observe_op(N, op, y, observed + 2 * (s / observe_steps));
        if (s % ck == 0 && s / ck < ncp)
            memcpy(cps + (s / ck) * size, y, sizeof(double) * (size_t)size);
    }

    /* adjoint source of row r: conj(w_r) O^dag on the top ADO */
#define ADD_SOURCE(r)                                                                    \
    do {                                                                                 \
        const double wr = weights[2 * (r)], wi = weights[2 * (r) + 1];                   \
        for (int i_ = 0; i_ < N; ++i_)                                                   \
            for (int j_ = 0; j_ < N; ++j_) {                                             \
                const double orr = op[(j_ * N + i_) * 2], oi = op[(j_ * N + i_) * 2 + 1]; \
                lam[(i_ * N + j_) * 2] += wr * orr - wi * oi;                            \
                lam[(i_ * N + j_) * 2 + 1] -= wr * oi + wi * orr;                        \
            }                                                                            \
    } while (0)

    memset(lam, 0, sizeof(double) * (size_t)size);
    ADD_SOURCE(rows - 1);
    for (int64_t c = (last - 1) / ck; last > 0 && c >= 0; --c) {
        const int64_t start = c * ck;
        const int64_t end = start + ck < last ? start + ck : last;
        memcpy(seg, cps + c * size, sizeof(double) * (size_t)size);
        for (int64_t n = start + 1; n < end; ++n) {
            memcpy(seg + (n - start) * size, seg + (n - start - 1) * size, sizeof(double) * (size_t)size);
            oracle_rk4_step(inst, h, seg + (n - start) * size, work);
        }
        for (int64_t n = end - 1; n >= start; --n) {
            const double* yn = seg + (n - start) * size;
            /* forward stages of step n */
            oracle_rhs(inst, yn, k);
            axpy_into(size, y2, yn, 0.5 * h, k);
            oracle_rhs(inst, y2, k);
            axpy_into(size, y3, yn, 0.5 * h, k);
            oracle_rhs(inst, y3, k);
            axpy_into(size, y4, yn, h, k);
            /* adjoint stages */
            oracle_rhs_adjoint(inst, lam, f);
            memcpy(acc, f, sizeof(double) * (size_t)size);
            axpy_into(size, a2, lam, 0.5 * h, f);
            oracle_rhs_adjoint(inst, a2, f);
            axpy_into(size, acc, acc, 2.0, f);
            axpy_into(size, a3, lam, 0.5 * h, f);
            oracle_rhs_adjoint(inst, a3, f);
            axpy_into(size, acc, acc, 2.0, f);
            axpy_into(size, a4, lam, h, f);
            oracle_rhs_adjoint(inst, a4, f);
            axpy_into(size, acc, acc, 1.0, f);
            oracle_adjoint_contract(inst, a4, yn, h / 6.0, g_h, g_pss, g_same, g_circ, g_cbk);
            oracle_adjoint_contract(inst, a3, y2, h / 3.0, g_h, g_pss, g_same, g_circ, g_cbk);
            oracle_adjoint_contract(inst, a2, y3, h / 3.0, g_h, g_pss, g_same, g_circ, g_cbk);
            oracle_adjoint_contract(inst, lam, y4, h / 6.0, g_h, g_pss, g_same, g_circ, g_cbk);
            axpy_into(size, lam, lam, h / 6.0, acc);
            if (n % observe_steps == 0)
                ADD_SOURCE(n / observe_steps);
        }
    }
#undef ADD_SOURCE
    if (g_state)
        memcpy(g_state, lam, sizeof(double) * (size_t)size);
    free(cps);
    free(seg);
    free(wk);
    return 0;
}

/* ======================================================================== */
/* time-dependent Hamiltonian (include/heom/e_field.hpp:64-87)              */
/* ======================================================================== */
/* The reference updates H before every RHS evaluation with the field at that
 * evaluation's time (a pre-evaluate action of the ODE, include/heom/ode.hpp:57-59); for
 * classical RK4 the stage times are t, t + h/2, t + h/2, t + h.  Here the four stage
 * Hamiltonians of each step are given: H4[s * 2NN .. ] for stage s. */
void oracle_rk4_step_timedep(const oracle_instance* inst, double h, double* y, double* work, const double* H4)
{
    const int64_t size = inst->matrices * inst->states * inst->states * 2;
    const int64_t nn2 = (int64_t)inst->states * inst->states * 2;
    oracle_instance st = *inst;
    double* k1 = work;
    double* k2 = work + size;
    double* k3 = work + 2 * size;
    double* k4 = work + 3 * size;
    double* t = work + 4 * size;
    const double* ks[4] = {k1, k2, k3, k4};
    const double a2[1] = {0.5};
    const double a3[2] = {0.0, 0.5};
    const double a4[3] = {0.0, 0.0, 1.0};
    const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

    st.hamiltonian = H4;
    oracle_rhs(&st, y, k1);
    oracle_rk_weighted_add(size, 1, h, t, y, a2, ks);
    st.hamiltonian = H4 + nn2;
    oracle_rhs(&st, t, k2);
    oracle_rk_weighted_add(size, 2, h, t, y, a3, ks);
    st.hamiltonian = H4 + 2 * nn2;
    oracle_rhs(&st, t, k3);
    oracle_rk_weighted_add(size, 3, h, t, y, a4, ks);
    st.hamiltonian = H4 + 3 * nn2;
    oracle_rhs(&st, t, k4);
    oracle_rk_weighted_add(size, 4, h, t, y, b, ks);
    memcpy(y, t, sizeof(double) * (size_t)size);
}

int64_t oracle_propagate_timedep(const oracle_instance* inst, double h, int64_t nsteps, int64_t observe_steps,
                                 double* y, double* work, const double* stage_h, double* tops)
{
    const int64_t top = (int64_t)inst->states * inst->states * 2;
    int64_t rows = 0;
    if (tops)
        memcpy(tops + (rows++) * top, y, sizeof(double) * (size_t)top);
    for (int64_t s = 1; s <= nsteps; ++s) {
        oracle_rk4_step_timedep(inst, h, y, work, stage_h + (s - 1) * 4 * top);
        if (tops && observe_steps > 0 && s % observe_steps == 0)
            memcpy(tops + (rows++) * top, y, sizeof(double) * (size_t)top);
    }
    return rows;
}

/* ======================================================================== */
/* thermal state search (include/heom/thermal_state_search_solver.hpp:267-299) */
/* ======================================================================== */

/* sigma <- sigma mx for every ADO (cl/heom/hierarchy_arithmetic_mmult_right.cl:15-50) */
void oracle_mmult_right(int states, int64_t matrices, const double* mx, double* sigma)
{
    const int N = states;
#pragma omp parallel if (matrices * N * N > 16384)
    {
        double* tmp = (double*)malloc(sizeof(double) * (size_t)(N * N * 2));
#pragma omp for schedule(static)
        for (int64_t id = 0; id < matrices; ++id) {
            double* a = sigma + id * N * N * 2;
            for (int i = 0; i < N; ++i) {
                for (int j = 0; j < N; ++j) {
                    double vr = 0.0, vi = 0.0;
                    for (int k = 0; k < N; ++k) {
                        const double ar = a[(i * N + k) * 2], ai = a[(i * N + k) * 2 + 1];
                        const double mr = mx[(k * N + j) * 2], mi = mx[(k * N + j) * 2 + 1];
                        vr = vr + (ar * mr - ai * mi);
                        vi = vi + (ar * mi + ai * mr);
                    }
                    tmp[(i * N + j) * 2] = vr;
                    tmp[(i * N + j) * 2 + 1] = vi;
                }
            }
            memcpy(a, tmp, sizeof(double) * (size_t)(N * N * 2));
        }
        free(tmp);
    }
}

/* One iteration of the search, in the reference's order: k = f_TSS(y) with the
 * pseudo-Hamiltonian in inst->hamiltonian (formulas (17)-(20)); k <- C k C^T;
 * y_u = k_u / (i (E_i - E_j) / hbar + eps - S_u) elementwise
 * (cl/heom/thermal_state_search.cl:36-48, S_u = pf_same_sum); y <- C^T y C.
 * c [N*N*2], eigenvalues [N*2], work [M*N*N*2]. */
void oracle_tss_step(const oracle_instance* inst, const double* c, const double* eigenvalues, double eps_r,
                     double eps_i, double* y, double* work)
{
    const int N = inst->states;
    const int64_t M = inst->matrices;
    double* ct = (double*)malloc(sizeof(double) * (size_t)(N * N * 2));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            ct[(i * N + j) * 2] = c[(j * N + i) * 2];
            ct[(i * N + j) * 2 + 1] = c[(j * N + i) * 2 + 1];
        }
    oracle_rhs_tss(inst, y, work);
    oracle_mmult_left(N, M, c, work);
    oracle_mmult_right(N, M, ct, work);
#pragma omp parallel for schedule(static) if (M * N * N > 16384)
    for (int64_t id = 0; id < M; ++id) {
        const double pr = inst->pf_same_sum[2 * id], pi = inst->pf_same_sum[2 * id + 1];
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                /* (i / hbar) (E_i - E_j) + (eps - S_u) */
                const double der = eigenvalues[2 * i] - eigenvalues[2 * j];
                const double dei = eigenvalues[2 * i + 1] - eigenvalues[2 * j + 1];
                const double dr = -dei / ORACLE_HBAR + (eps_r - pr);
                const double di = der / ORACLE_HBAR + (eps_i - pi);
                const int64_t e = (id * N * N + i * N + j) * 2;
                const double ar = work[e], ai = work[e + 1];
                const double den = dr * dr + di * di;
                y[e] = (ar * dr + ai * di) / den;
                y[e + 1] = (ai * dr - ar * di) / den;
            }
    }
    oracle_mmult_left(N, M, ct, y);
    oracle_mmult_right(N, M, c, y);
    free(ct);
}

