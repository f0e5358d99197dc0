// Argument-type extraction for user lambdas (the role of the reference's
// detail::FunctionTraits, include/rxmesh/util/meta.h): the query dispatcher
// deduces the handle and iterator types from the lambda's parameters.
#pragma once
#include <cstddef>
#include <tuple>
#include <type_traits>

namespace rxmesh {
namespace detail {

template <typename T>
struct FunctionTraits : public FunctionTraits<decltype(&T::operator())>
{
};

template <typename ClassT, typename RetT, typename... ArgsT>
struct FunctionTraits<RetT (ClassT::*)(ArgsT...) const>
{
    static constexpr std::size_t arity = sizeof...(ArgsT);
    using result_type                  = RetT;
    template <std::size_t i>
    struct arg
    {
        using type = std::remove_cv_t<std::remove_reference_t<std::tuple_element_t<i, std::tuple<ArgsT...>>>>;
    };
};

template <typename ClassT, typename RetT, typename... ArgsT>
struct FunctionTraits<RetT (ClassT::*)(ArgsT...)> : FunctionTraits<RetT (ClassT::*)(ArgsT...) const>
{
};

template <typename RetT, typename... ArgsT>
struct FunctionTraits<RetT (*)(ArgsT...)>
{
    static constexpr std::size_t arity = sizeof...(ArgsT);
    using result_type                  = RetT;
    template <std::size_t i>
    struct arg
    {
        using type = std::remove_cv_t<std::remove_reference_t<std::tuple_element_t<i, std::tuple<ArgsT...>>>>;
    };
};

}  // namespace detail
}  // namespace rxmesh
