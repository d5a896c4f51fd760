// cmbl_function.hpp - the function-with-gradient interfaces of Cosmo++, for
// builds that do not have the Cosmo++ headers.
//
// The declarations are interface-identical to the reference's
//   Math::FunctionMultiDim / FunctionMultiToMulti / RealFunctionMultiDim /
//   RealFunctionMultiToMulti      (include/function.hpp:111-179)
//   Math::LikelihoodFunction / LikelihoodWithDerivs
//                                 (include/likelihood_function.hpp:7-47)
// and use the same include guards, so that a translation unit which also
// includes the Cosmo++ headers sees exactly one definition: whichever header
// comes first provides the types, and classes derived from them (e.g.
// cmbl::MapLikelihood) plug into Cosmo++'s LBFGS, HMC and CG unchanged.
#ifndef COSMO_PP_FUNCTION_HPP
#define COSMO_PP_FUNCTION_HPP

#include <vector>

namespace Math
{

/// A function of one variable.
template <typename VarType, typename ValType>
class Function
{
public:
    typedef VarType VariableType;
    typedef ValType ValueType;

    Function() {}
    virtual ~Function() {}

    virtual ValueType evaluate(VariableType x) const = 0;
};

/// A function of two variables.
template <typename Var1Type, typename Var2Type, typename ValType>
class Function2
{
public:
    typedef Var1Type Variable1Type;
    typedef Var2Type Variable2Type;
    typedef ValType ValueType;

    Function2() {}
    virtual ~Function2() {}

    virtual ValueType evaluate(Variable1Type x1, Variable2Type x2) const = 0;
};

/// A function of three variables.
template <typename Var1Type, typename Var2Type, typename Var3Type, typename ValType>
class Function3
{
public:
    typedef Var1Type Variable1Type;
    typedef Var2Type Variable2Type;
    typedef Var3Type Variable3Type;
    typedef ValType ValueType;

    Function3() {}
    virtual ~Function3() {}

    virtual ValueType evaluate(Variable1Type x1, Variable2Type x2, Variable3Type x3) const = 0;
};

typedef Function<double, double> RealFunction;

/// A function of many variables returning one value (-2lnL for a likelihood).
template <typename VarType, typename ValType>
class FunctionMultiDim
{
public:
    typedef VarType VariableType;
    typedef ValType ValueType;

    FunctionMultiDim() {}
    virtual ~FunctionMultiDim() {}

    virtual ValueType evaluate(const std::vector<VariableType>& x) const = 0;
};

/// A function of many variables returning many values (a gradient); res is resized by the callee.
template <typename VarType, typename ValType>
class FunctionMultiToMulti
{
public:
    typedef VarType VariableType;
    typedef ValType ValueType;

    FunctionMultiToMulti() {}
    virtual ~FunctionMultiToMulti() {}

    virtual void evaluate(const std::vector<VariableType>& x, std::vector<ValueType>* res) const = 0;
};

typedef FunctionMultiDim<double, double> RealFunctionMultiDim;
typedef FunctionMultiToMulti<double, double> RealFunctionMultiToMulti;

}  // namespace Math

#endif

#ifndef COSMO_PP_LIKELIHOOD_FUNCTION_HPP
#define COSMO_PP_LIKELIHOOD_FUNCTION_HPP

namespace Math
{

/// -2ln(likelihood) of a parameter vector.
class LikelihoodFunction
{
public:
    virtual ~LikelihoodFunction() {}
    virtual double calculate(double* params, int nParams) = 0;
    virtual double calculateExact(double* params, int nParams) { return calculate(params, nParams); }
};

/// -2ln(likelihood) with its partial derivatives, one component per call.
class LikelihoodWithDerivs : public LikelihoodFunction
{
public:
    virtual ~LikelihoodWithDerivs() {}
    virtual double calculate(double* params, int nParams) = 0;
    virtual double calculateDeriv(double* params, int nParams, int i) = 0;
};

}  // namespace Math

#endif
