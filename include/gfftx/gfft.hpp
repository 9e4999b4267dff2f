// gfft.hpp -- GFFT's plan-and-call surface with a GPU parallelisation policy HIP<Dev>.
//
// Host-only C++17 over the C ABI in gfftx.h; it needs no HIP headers and links against
// libgfftx.so.  Names, IDs and defaults follow the reference:
//   value types   DOUBLE/FLOAT/COMPLEX_DOUBLE/COMPLEX_FLOAT      src/gfftpolicy.h:78-137
//   transforms    DFT/IDFT/RDFT/IRDFT                            src/gfftpolicy.h:225-297
//   places        IN_PLACE/OUT_OF_PLACE                          src/gfftpolicy.h:143-215
//   interfaces    AbstractFFT_inp<T>::fft(T*),
//                 AbstractFFT_oop<T>::fft(const T*, T*)          src/gfftpolicy.h:46-58
//   Transform<N,VType,Type,Dim,Parall,Place>::Instance           src/gfftgen.h:46-87
//   GenerateTransform<NList,VType,Type,Dim,Parall,Place>
//       ::CreateTransformObject(n, vtype, trans, dim, parall, place)  src/gfftgen.h:189-233
//   group defaults DFT / OUT_OF_PLACE                            src/gfftparamgroups.h:44-78
// HIP<Dev> replaces Serial/OpenMP<NT>: the transform runs on GPU `Dev`.  Its ID is
// HIP_ID_BASE + Dev so that it never collides with OpenMP<NT>::ID = NT-1.
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
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../gfftx.h"

namespace gfftx {

typedef unsigned int id_t;

template <unsigned long N> struct ulong_ {
    static const unsigned long value = N;
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
    static const int Device = Dev;
    static const char* name() { return "hip"; }
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

template <class VType, class Type, class Place, class Parall> struct Exec;

template <class VType, class Type, class Parall>
struct Exec<VType, Type, IN_PLACE, Parall> : public AbstractFFT_inp<typename VType::ValueType> {
    typedef typename VType::ValueType T;
    PlanHolder h;
    explicit Exec(unsigned long n) : h(n, VType::ID, Type::ID, IN_PLACE::ID, Parall::Device) {}
    void fft(T* data) override
    {
        check(gfftx_exec_inplace(h.get(), data, nullptr), "fft(T*)");
        check(gfftx_stream_synchronize(nullptr), "fft(T*)");
    }
};

template <class VType, class Type, class Parall>
struct Exec<VType, Type, OUT_OF_PLACE, Parall> : public AbstractFFT_oop<typename VType::ValueType> {
    typedef typename VType::ValueType T;
    PlanHolder h;
    explicit Exec(unsigned long n) : h(n, VType::ID, Type::ID, OUT_OF_PLACE::ID, Parall::Device) {}
    void fft(const T* src, T* dst) override
    {
        check(gfftx_exec_oop(h.get(), src, dst, nullptr), "fft(const T*, T*)");
        check(gfftx_stream_synchronize(nullptr), "fft(const T*, T*)");
    }
};

}  // namespace detail

// Transform<N, VType, Type, Dim, Parall, Place>: one fixed transform (src/gfftgen.h:46-87).
// N is ulong_<n>; n is the complex length (half the real length for RDFT/IRDFT).
template <class N, class VType, class Type = DFT, class Dim = ulong_<1>, class Parall = HIP<0>,
          class Place = OUT_OF_PLACE>
class Transform {
    static_assert(Dim::value == 1, "only one-dimensional transforms exist (src/gfftgen.h:154)");

public:
    typedef VType ValueType;
    typedef Type TransformType;
    typedef Parall ParallType;
    typedef Place PlaceType;
    static const unsigned long Len = N::value;
    typedef typename Place::template Interface<typename VType::ValueType>::Result ReturnType;

    class Instance : public detail::Exec<VType, Type, Place, Parall> {
    public:
        Instance() : detail::Exec<VType, Type, Place, Parall>(N::value) {}
    };

    static ReturnType* Create() { return new Instance(); }
};

// Parameter lists (the typelists of src/gfftparamgroups.h as variadic packs).
template <class... Ts> struct List {};
template <unsigned long... Ns> struct NList {};

// GenPowerList<Min, Max, Base>: lengths Base^Min .. Base^Max (src/typelistgen.h:116-147).
namespace detail {
template <unsigned long B, unsigned long P> struct Pow {
    static const unsigned long value = B * Pow<B, P - 1>::value;
};
template <unsigned long B> struct Pow<B, 0> {
    static const unsigned long value = 1;
};
template <unsigned long Min, unsigned long Max, unsigned long B, class Seq> struct PowSeq;
template <unsigned long Min, unsigned long Max, unsigned long B, std::size_t... I>
struct PowSeq<Min, Max, B, std::index_sequence<I...>> {
    typedef NList<Pow<B, Min + I>::value...> type;
};
}  // namespace detail
template <unsigned long Min, unsigned long Max, unsigned long Base = 2>
using GenPowerList = typename detail::PowSeq<Min, Max, Base, std::make_index_sequence<Max - Min + 1>>::type;

namespace detail {
template <class X> struct AsList {
    typedef List<X> type;
};
template <class... Ts> struct AsList<List<Ts...>> {
    typedef List<Ts...> type;
};
}  // namespace detail

// GenerateTransform: registers every combination of the parameter lists and creates
// objects by run-time IDs (src/gfftgen.h:189-233).  TransType/Parall/Place may be single
// classes or List<...>; all places must share one interface, as in the reference.
template <class NL, class VType, class TransType = DFT, class Dim = ulong_<1>, class Parall = HIP<0>,
          class Place = OUT_OF_PLACE>
class GenerateTransform;

template <unsigned long... Ns, class VType, class TransType, class Dim, class Parall, class Place>
class GenerateTransform<NList<Ns...>, VType, TransType, Dim, Parall, Place> {
    struct Key {
        unsigned long n;
        id_t vtype, trans, dim, parall, place;
    };
    std::vector<Key> keys_;

    template <class T, class P, class Pl> void add_one()
    {
        for (unsigned long n : {Ns...}) keys_.push_back(Key{n, VType::ID, T::ID, (id_t)Dim::value, P::ID, Pl::ID});
    }
    template <class T, class P, class... Pls> void add_places(List<Pls...>)
    {
        (add_one<T, P, Pls>(), ...);
    }
    template <class T, class... Ps> void add_paralls(List<Ps...>)
    {
        (add_places<T, Ps>(typename detail::AsList<Place>::type()), ...);
    }
    template <class... Ts> void add_types(List<Ts...>)
    {
        (add_paralls<Ts>(typename detail::AsList<Parall>::type()), ...);
    }
    template <class X> struct FirstOf {
        typedef X type;
    };
    template <class X, class... Rest> struct FirstOf<List<X, Rest...>> {
        typedef X type;
    };
    template <class P, class... Ps> static int device_of(id_t parall_id, List<P, Ps...>)
    {
        if (P::ID == parall_id) return P::Device;
        if constexpr (sizeof...(Ps) > 0) return device_of(parall_id, List<Ps...>());
        return -1;
    }

public:
    typedef typename FirstOf<Place>::type PlaceType;
    typedef typename PlaceType::template Interface<typename VType::ValueType>::Result ObjectType;

    GenerateTransform() { add_types(typename detail::AsList<TransType>::type()); }

    ObjectType* CreateTransformObject(unsigned long n, unsigned long vtype_id, unsigned long trans_id = DFT::ID,
                                      unsigned long dim = 1,
                                      unsigned long parall_id = FirstOf<Parall>::type::ID,
                                      unsigned long place_id = PlaceType::ID)
    {
        bool found = false;
        for (const Key& k : keys_)
            found = found || (k.n == n && k.vtype == vtype_id && k.trans == trans_id && k.dim == dim &&
                              k.parall == parall_id && k.place == place_id);
        if (!found) throw TransformFactoryError::Exception();
        const int dev = device_of((id_t)parall_id, typename detail::AsList<Parall>::type());
        return make(n, (id_t)trans_id, dev);
    }

private:
    ObjectType* make(unsigned long n, id_t trans, int dev)
    {
        // The object type is fixed by the interface; the transform type is a run-time ID.
        return new RuntimeExec(n, trans, dev);
    }

    class RuntimeExec;
};

namespace detail {
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
}  // namespace detail

template <unsigned long... Ns, class VType, class TransType, class Dim, class Parall, class Place>
class GenerateTransform<NList<Ns...>, VType, TransType, Dim, Parall, Place>::RuntimeExec
    : public detail::RuntimeObject<VType, typename GenerateTransform<NList<Ns...>, VType, TransType, Dim, Parall,
                                                                       Place>::PlaceType> {
public:
    RuntimeExec(unsigned long n, id_t trans, int dev)
        : detail::RuntimeObject<VType, typename GenerateTransform::PlaceType>(n, trans, dev)
    {
    }
};

}  // namespace gfftx
