// gfft.hpp -- GFFT's plan-and-call surface with a GPU parallelisation policy HIP<Dev>.
//
// Host-only C++17 over the C ABI in gfftx.h; it needs no HIP headers and links against
// libgfftx.so.  Names, IDs and defaults follow the reference, so that GFFT user code compiles
// against this header (namespace GFFT is an alias of gfftx):
//   value types   DOUBLE/FLOAT/COMPLEX_DOUBLE/COMPLEX_FLOAT      src/gfftpolicy.h:78-137
//   transforms    DFT/IDFT/RDFT/IRDFT                            src/gfftpolicy.h:225-297
//   places        IN_PLACE/OUT_OF_PLACE                          src/gfftpolicy.h:143-215
//   parallel      Serial, OpenMP<NT>                             src/gfftpolicy.h:371-453
//   interfaces    AbstractFFT_inp<T>::fft(T*),
//                 AbstractFFT_oop<T>::fft(const T*, T*)          src/gfftpolicy.h:46-58
//   Transform<N,VType,Type,Dim,Parall,Place,IDN>::Instance, ::ID src/gfftgen.h:46-87
//   GenerateTransform<NList,VType,Type,Dim,Parall,Place>
//       ::CreateTransformObject(n, vtype, trans, dim, parall, place), ::Result
//                                                                src/gfftgen.h:189-233
//   groups        ValueTypeGroup, TransformTypeGroup,
//                 ParallelizationGroup, PlaceGroup               src/gfftparamgroups.h:35-78
//   length lists  TYPELIST_n, GenNumList, GenPowerList           loki/Typelist.h:33-,
//                                                                src/typelistgen.h:101-147
// Parameter lists are Loki typelists, as in the reference: when Loki's Typelist.h is not
// included first, this header declares the subset GFFT's interface uses (Typelist,
// NullType, TYPELIST_1..TYPELIST_8, TypeTraits<T>::isStdFundamental).
// HIP<Dev> runs the transform on GPU `Dev`; its ID is HIP_ID_BASE + Dev so that it never
// collides with OpenMP<NT>::ID = NT-1.  Serial and OpenMP<NT> keep their IDs and run on GPU 0:
// the engine has no CPU path, and code written for them runs unchanged.
// fft() is synchronous, as GFFT's is: it returns once the result is in `data`/`dst`
// (host or device memory).
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../gfftx.h"

// --- the part of Loki's interface GFFT's parameter lists use (loki/Typelist.h,
// loki/NullType.h, loki/TypeTraits.h): skipped where Loki's own headers came first
#if !defined(TYPELIST_INC_) && !defined(GFFTX_LOKI_TYPELIST)
#define GFFTX_LOKI_TYPELIST
namespace Loki {
class NullType {};
template <class T, class U> struct Typelist {
    typedef T Head;
    typedef U Tail;
};
}  // namespace Loki
#define TYPELIST_1(T1) ::Loki::Typelist<T1, ::Loki::NullType>
#define TYPELIST_2(T1, T2) ::Loki::Typelist<T1, TYPELIST_1(T2)>
#define TYPELIST_3(T1, T2, T3) ::Loki::Typelist<T1, TYPELIST_2(T2, T3)>
#define TYPELIST_4(T1, T2, T3, T4) ::Loki::Typelist<T1, TYPELIST_3(T2, T3, T4)>
#define TYPELIST_5(T1, T2, T3, T4, T5) ::Loki::Typelist<T1, TYPELIST_4(T2, T3, T4, T5)>
#define TYPELIST_6(T1, T2, T3, T4, T5, T6) ::Loki::Typelist<T1, TYPELIST_5(T2, T3, T4, T5, T6)>
#define TYPELIST_7(T1, T2, T3, T4, T5, T6, T7) ::Loki::Typelist<T1, TYPELIST_6(T2, T3, T4, T5, T6, T7)>
#define TYPELIST_8(T1, T2, T3, T4, T5, T6, T7, T8) ::Loki::Typelist<T1, TYPELIST_7(T2, T3, T4, T5, T6, T7, T8)>
#endif
#if !defined(TYPETRAITS_INC_) && !defined(GFFTX_LOKI_TYPETRAITS)
#define GFFTX_LOKI_TYPETRAITS
namespace Loki {
template <typename T> struct TypeTraits {
    enum { isStdFundamental = std::is_fundamental<T>::value };
};
}  // namespace Loki
#endif

namespace gfftx {

typedef unsigned int id_t;
typedef long long long_t;
typedef unsigned long ulong_t;

// Length holders: value N, ID N-1 (src/typelistgen.h:32-35)
template <ulong_t N> struct ulong_ {
    static const ulong_t value = N;
    static const id_t ID = (id_t)(N - 1);
};
template <long_t N> struct long_ {
    static const long_t value = N;
    static const id_t ID = (id_t)(N - 1);
};

struct DOUBLE {
    static const id_t ID = 0;
    typedef double base_type;
    typedef double ValueType;
    static const char* name() { return "double"; }
};
struct FLOAT {
    static const id_t ID = 1;
    typedef float base_type;
    typedef float ValueType;
    static const char* name() { return "float"; }
};
struct COMPLEX_DOUBLE {
    static const id_t ID = 2;
    typedef double base_type;
    typedef std::complex<double> ValueType;
    static const char* name() { return "std::complex<double>"; }
};
struct COMPLEX_FLOAT {
    static const id_t ID = 3;
    typedef float base_type;
    typedef std::complex<float> ValueType;
    static const char* name() { return "std::complex<float>"; }
};

struct IDFT;
struct IRDFT;
struct DFT {
    static const id_t ID = 0;
    static const int Sign = 1;
    typedef IDFT Inverse;
};
struct IDFT {
    static const id_t ID = 1;
    static const int Sign = -1;
    typedef DFT Inverse;
};
struct RDFT {
    static const id_t ID = 2;
    static const int Sign = 1;
    typedef IRDFT Inverse;
};
struct IRDFT {
    static const id_t ID = 3;
    static const int Sign = -1;
    typedef RDFT Inverse;
};

template <typename T> class AbstractFFT_inp {
public:
    virtual void fft(T*) = 0;
    virtual ~AbstractFFT_inp() {}
};

template <typename T> class AbstractFFT_oop {
public:
    virtual void fft(const T*, T*) = 0;
    virtual ~AbstractFFT_oop() {}
};

struct IN_PLACE {
    static const id_t ID = 0;
    template <class T> struct Interface {
        typedef AbstractFFT_inp<T> Result;
    };
    static const char* name() { return "in-place"; }
};
struct OUT_OF_PLACE {
    static const id_t ID = 1;
    template <class T> struct Interface {
        typedef AbstractFFT_oop<T> Result;
    };
    static const char* name() { return "out-of-place"; }
};

static const id_t HIP_ID_BASE = 64;

// GPU parallelisation policy: the transform runs on HIP device Dev.
template <int Dev = 0> struct HIP {
    static const id_t ID = HIP_ID_BASE + (id_t)Dev;
    static const ulong_t NParProc = 1;
    static const int Device = Dev;
    static const char* name() { return "hip"; }
};

// GFFT's CPU policies, kept with their IDs (Serial 0, OpenMP<NT> NT-1; OpenMP<0> and
// OpenMP<1> are Serial, src/gfftpolicy.h:371-453).  Here they run on GPU 0.
struct Serial {
    static const id_t ID = 0;
    static const ulong_t NParProc = 1;
    static const int Device = 0;
    static const char* name() { return "serial (GPU 0)"; }
};
template <unsigned int NT> struct OpenMP {
    static const id_t ID = NT - 1;
    static const ulong_t NParProc = NT;
    static const int Device = 0;
    static const char* name() { return "openmp (GPU 0)"; }
};
template <> struct OpenMP<0> : public Serial {};
template <> struct OpenMP<1> : public Serial {};

// Parameter groups (src/gfftparamgroups.h:35-78): FullList, Length and Default as the
// reference states them (ParallelizationGroup's Length reads 3 there for its four entries).
struct ValueTypeGroup {
    typedef TYPELIST_4(DOUBLE, FLOAT, COMPLEX_DOUBLE, COMPLEX_FLOAT) FullList;
    static const ulong_t Length = 4;
    typedef DOUBLE Default;
};
struct TransformTypeGroup {
    typedef TYPELIST_4(DFT, IDFT, RDFT, IRDFT) FullList;
    static const ulong_t Length = 4;
    typedef DFT Default;
};
struct ParallelizationGroup {
    typedef TYPELIST_4(OpenMP<1>, OpenMP<2>, OpenMP<4>, OpenMP<8>) FullList;
    static const ulong_t Length = 3;
    typedef Serial Default;
};
struct PlaceGroup {
    typedef TYPELIST_2(IN_PLACE, OUT_OF_PLACE) FullList;
    static const ulong_t Length = 2;
    typedef OUT_OF_PLACE Default;
};

// Error type of an unknown combination, as TransformFactoryError (src/gfftgen.h:133-145).
struct TransformFactoryError {
    struct Exception : public std::exception {
        const char* what() const throw()
        {
            return "Requested transform is not compiled! Check your instantiation of GenerateTransform class!";
        }
    };
};

// Any other failure of the engine (device missing, allocation, launch).
struct Error : public std::runtime_error {
    int status;
    Error(int s, const std::string& where)
        : std::runtime_error(where + ": " + gfftx_strerror(s) + " " + gfftx_last_error()), status(s)
    {
    }
};

inline void check(int s, const char* where)
{
    if (s == GFFTX_SUCCESS) return;
    if (s == GFFTX_ERR_UNSUPPORTED) throw TransformFactoryError::Exception();
    throw Error(s, where);
}

namespace detail {

// Owns one gfftx plan; created on first use so that a Transform object can be declared
// before a device is selected, as GFFT objects are declared freely.
class PlanHolder {
public:
    PlanHolder(unsigned long n, id_t vtype, id_t ttype, id_t place, int dev)
        : n_(n), vtype_(vtype), ttype_(ttype), place_(place), dev_(dev)
    {
    }
    ~PlanHolder()
    {
        if (plan_) gfftx_plan_destroy(plan_);
    }
    PlanHolder(const PlanHolder&) = delete;
    PlanHolder& operator=(const PlanHolder&) = delete;
    gfftx_plan get()
    {
        if (!plan_) check(gfftx_plan_create(&plan_, n_, (int)vtype_, (int)ttype_, (int)place_, dev_, 0u),
                          "gfftx_plan_create");
        return plan_;
    }

private:
    unsigned long n_;
    id_t vtype_, ttype_, place_;
    int dev_;
    gfftx_plan plan_ = nullptr;
};

// The transform object for a place: the interface GFFT's place policy names, with fft()
// calling the C ABI and waiting for the result.
template <class VType, class Place> class RuntimeObject;
template <class VType>
class RuntimeObject<VType, IN_PLACE> : public AbstractFFT_inp<typename VType::ValueType> {
public:
    RuntimeObject(unsigned long n, id_t trans, int dev) : h_(n, VType::ID, trans, IN_PLACE::ID, dev) {}
    void fft(typename VType::ValueType* data) override
    {
        check(gfftx_exec_inplace(h_.get(), data, nullptr), "fft(T*)");
        check(gfftx_stream_synchronize(nullptr), "fft(T*)");
    }

private:
    PlanHolder h_;
};
template <class VType>
class RuntimeObject<VType, OUT_OF_PLACE> : public AbstractFFT_oop<typename VType::ValueType> {
public:
    RuntimeObject(unsigned long n, id_t trans, int dev) : h_(n, VType::ID, trans, OUT_OF_PLACE::ID, dev) {}
    void fft(const typename VType::ValueType* src, typename VType::ValueType* dst) override
    {
        check(gfftx_exec_oop(h_.get(), src, dst, nullptr), "fft(const T*, T*)");
        check(gfftx_stream_synchronize(nullptr), "fft(const T*, T*)");
    }

private:
    PlanHolder h_;
};

// --- typelist helpers: a single class counts as a one-element list
template <class X> struct AsList {
    typedef Loki::Typelist<X, Loki::NullType> type;
};
template <class H, class T> struct AsList<Loki::Typelist<H, T>> {
    typedef Loki::Typelist<H, T> type;
};
template <> struct AsList<Loki::NullType> {
    typedef Loki::NullType type;
};
template <class A, class B> struct Append;
template <class B> struct Append<Loki::NullType, B> {
    typedef B type;
};
template <class H, class T, class B> struct Append<Loki::Typelist<H, T>, B> {
    typedef Loki::Typelist<H, typename Append<T, B>::type> type;
};
template <class L> struct Head {
    typedef typename AsList<L>::type::Head type;
};
// every place in the list has the first one's interface
template <class L, class P> struct SameInterface {
    static const bool value = true;
};
template <class H, class T, class P> struct SameInterface<Loki::Typelist<H, T>, P> {
    static const bool value = std::is_same<typename H::template Interface<int>::Result,
                                           typename P::template Interface<int>::Result>::value &&
                              SameInterface<T, P>::value;
};

}  // namespace detail

// Transform<N, VType, Type, Dim, Parall, Place, IDN>: one fixed transform (src/gfftgen.h:46-87).
// N is a length holder (ulong_<n>); n is the complex length (half the real length for
// RDFT/IRDFT).  IDN is the factory ID GFFT gives the combination (N::ID by default).
template <class N, class VType, class Type = DFT, class Dim = ulong_<1>, class Parall = HIP<0>,
          class Place = OUT_OF_PLACE, id_t IDN = N::ID>
class Transform {
    static_assert(Dim::value == 1, "only one-dimensional transforms exist (src/gfftgen.h:154)");

public:
    typedef VType ValueType;
    typedef Type TransformType;
    typedef Parall ParallType;
    typedef Place PlaceType;
    typedef typename Place::template Interface<typename VType::ValueType>::Result ReturnType;

    enum { ID = IDN };
    static const long_t Len = (long_t)N::value;

    class Instance : public detail::RuntimeObject<VType, Place> {
    public:
        Instance() : detail::RuntimeObject<VType, Place>((unsigned long)N::value, Type::ID, Parall::Device) {}
    };

    static ReturnType* Create() { return new Instance(); }
};

// Typelist of holders N = Begin..End (src/typelistgen.h:101-114)
template <long_t Begin, long_t End, template <long_t> class Holder = long_> struct GenNumList {
    typedef Loki::Typelist<Holder<Begin>, typename GenNumList<Begin + 1, End, Holder>::Result> Result;
};
template <long_t End, template <long_t> class Holder> struct GenNumList<End, End, Holder> {
    typedef Loki::Typelist<Holder<End>, Loki::NullType> Result;
};

// Length M^P with value and ID = value - 1 (src/typelistgen.h:117-131)
template <long_t M, long_t P> struct PowerHolder {
    static const long_t N = M * PowerHolder<M, P - 1>::N;
    static const long_t ID = N - 1;
    static const long_t value = N;
};
template <long_t M> struct PowerHolder<M, 0> {
    static const long_t N = 1;
    static const long_t ID = 0;
    static const long_t value = 1;
};

// Typelist of PowerHolder<M, P>, P = PowerBegin..PowerEnd (src/typelistgen.h:137-147)
template <long_t PowerBegin, long_t PowerEnd, long_t M> struct GenPowerList {
    typedef Loki::Typelist<PowerHolder<M, PowerBegin>, typename GenPowerList<PowerBegin + 1, PowerEnd, M>::Result>
        Result;
};
template <long_t PowerEnd, long_t M> struct GenPowerList<PowerEnd, PowerEnd, M> {
    typedef Loki::Typelist<PowerHolder<M, PowerEnd>, Loki::NullType> Result;
};

namespace detail {

// Every combination of the parameter lists as Transform classes: lengths outermost, then
// transform types, parallelisations and places.
template <class N, class V, class T, class Dim, class P, class PlL> struct ForPlaces {
    typedef Loki::NullType type;
};
template <class N, class V, class T, class Dim, class P, class Pl, class R>
struct ForPlaces<N, V, T, Dim, P, Loki::Typelist<Pl, R>> {
    typedef Loki::Typelist<Transform<N, V, T, Dim, P, Pl>, typename ForPlaces<N, V, T, Dim, P, R>::type> type;
};
template <class N, class V, class T, class Dim, class PL, class PlL> struct ForParalls {
    typedef Loki::NullType type;
};
template <class N, class V, class T, class Dim, class P, class R, class PlL>
struct ForParalls<N, V, T, Dim, Loki::Typelist<P, R>, PlL> {
    typedef typename Append<typename ForPlaces<N, V, T, Dim, P, PlL>::type,
                            typename ForParalls<N, V, T, Dim, R, PlL>::type>::type type;
};
template <class N, class V, class TL, class Dim, class PL, class PlL> struct ForTypes {
    typedef Loki::NullType type;
};
template <class N, class V, class T, class R, class Dim, class PL, class PlL>
struct ForTypes<N, V, Loki::Typelist<T, R>, Dim, PL, PlL> {
    typedef typename Append<typename ForParalls<N, V, T, Dim, PL, PlL>::type,
                            typename ForTypes<N, V, R, Dim, PL, PlL>::type>::type type;
};
template <class NL, class V, class TL, class Dim, class PL, class PlL> struct Product {
    typedef Loki::NullType type;
};
template <class N, class R, class V, class TL, class Dim, class PL, class PlL>
struct Product<Loki::Typelist<N, R>, V, TL, Dim, PL, PlL> {
    typedef typename Append<typename ForTypes<N, V, TL, Dim, PL, PlL>::type,
                            typename Product<R, V, TL, Dim, PL, PlL>::type>::type type;
};

// Run-time keys of every combination in a list of Transform classes.
struct Key {
    unsigned long n;
    id_t vtype, trans, dim, parall, place;
    int device;
};
template <class L> struct Keys {
    static void add(std::vector<Key>&) {}
};
template <class H, class T> struct Keys<Loki::Typelist<H, T>> {
    static void add(std::vector<Key>& v)
    {
        v.push_back(Key{(unsigned long)H::Len, H::ValueType::ID, H::TransformType::ID, 1u, H::ParallType::ID,
                        H::PlaceType::ID, H::ParallType::Device});
        Keys<T>::add(v);
    }
};

}  // namespace detail

// GenerateTransform: registers every combination of the parameter lists and creates objects
// by run-time IDs (src/gfftgen.h:189-233).  NList is a typelist of length holders;
// TransType, Parall and Place may be single classes or typelists; all places share one
// interface, as in the reference.  Result is the typelist of the Transform classes.
template <class NList, class VType, class TransType = TransformTypeGroup::Default, class Dim = ulong_<1>,
          class Parall = ParallelizationGroup::Default, class Place = PlaceGroup::Default>
class GenerateTransform {
    typedef typename detail::AsList<NList>::type NL;
    typedef typename detail::AsList<TransType>::type TL;
    typedef typename detail::AsList<Parall>::type PL;
    typedef typename detail::AsList<Place>::type PlL;
    std::vector<detail::Key> keys_;

public:
    typedef typename detail::Product<NL, VType, TL, Dim, PL, PlL>::type Result;
    typedef typename detail::Head<Place>::type PlaceType;
    typedef typename PlaceType::template Interface<typename VType::ValueType>::Result ObjectType;
    static_assert(detail::SameInterface<PlL, PlaceType>::value,
                  "the places of one GenerateTransform must share one interface (src/gfftgen.h:203)");

    GenerateTransform() { detail::Keys<Result>::add(keys_); }

    // The defaults for parall_id and place_id are the first entries of the Parall and Place
    // lists (the reference takes the group defaults, Serial and OUT_OF_PLACE, which coincide
    // when the lists are left at their defaults).
    ObjectType* CreateTransformObject(ulong_t n, ulong_t vtype_id, ulong_t trans_id = TransformTypeGroup::Default::ID,
                                      ulong_t dim = 1,
                                      ulong_t parall_id = detail::Head<Parall>::type::ID,
                                      ulong_t place_id = PlaceType::ID)
    {
        for (const detail::Key& k : keys_)
            if (k.n == n && k.vtype == vtype_id && k.trans == trans_id && k.dim == dim && k.parall == parall_id &&
                k.place == place_id)
                return new detail::RuntimeObject<VType, PlaceType>(n, (id_t)trans_id, k.device);
        throw TransformFactoryError::Exception();
    }
};

}  // namespace gfftx

// GFFT's namespace: user code that names GFFT:: or says `using namespace GFFT;` compiles
// against this header unchanged.
namespace GFFT = gfftx;
