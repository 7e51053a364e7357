classdef GpuOperator < rooftopMoM.Product
% GpuOperator: The Green's matrix of a Problem applied on the GPU, a drop-in for rooftopMoM.Operator
%   (src/+rooftopMoM/Operator.m:7-54). rooftopMoM.Solver accepts it as any Product; rooftopMoM.gpu.GpuSolver runs the
%   whole Krylov loop on the GPU with it.
%       operator = rooftopMoM.gpu.GpuOperator(problem, 'product', 'splitFFT', 'precision', 'mixed');

    properties (Access = private)
        handle_ uint64;
    end

    methods
        function obj = GpuOperator(problem, options)
        % GpuOperator: Creates the product.
        %
        %   Optional name-value inputs:
        %       'product': 'fft', 'splitFFT' or 'leanSplitFFT'; all three run the same GPU product (default
        %           'splitFFT', Operator.m:23-24).
        %       'modifier': @real or @imag applied to the generators (default none, Operator.m:25).
        %       'precision': 'c128', 'c64' or 'mixed' (both held, for mixed-precision solves; default 'c128').

            arguments
                problem (1, 1) rooftopMoM.Problem
                options.product (1, :) char {mustBeMember(options.product, {'fft', 'splitFFT', 'leanSplitFFT'})} = ...
                    'splitFFT'
                options.modifier function_handle = function_handle.empty
                options.precision (1, :) char {mustBeMember(options.precision, {'c128', 'c64', 'mixed'})} = 'c128'
            end

            kind = find(strcmp(options.product, {'fft', 'splitFFT', 'leanSplitFFT'})) - 1;
            precision = find(strcmp(options.precision, {'c128', 'c64', 'mixed'})) - 1;
            modifier = 0;

            if ~isempty(options.modifier)
                modifier = find(strcmp(func2str(options.modifier), {'real', 'imag'}));

                if isempty(modifier)
                    error('rooftopMoM:operator:invalidModifier', 'The GPU product takes the modifier @real or @imag.');
                end
            end

            problemHandle = rooftopMoM.gpu.problemHandle(problem);
            cleanup = onCleanup(@() rooftop_mex('free', 'problem', problemHandle));
            obj.handle_ = rooftop_mex('operator', problemHandle, kind, precision, modifier);
            obj.problem_ = problem;
            obj.applicator_ = rooftopMoM.gpu.GpuApplicator(obj.handle_, min(precision, 1));
        end

        function delete(obj)
            if ~isempty(obj.handle_)
                rooftop_mex('free', 'product', obj.handle_);
            end
        end

        function handle = handle(obj)
            handle = obj.handle_;
        end
    end
end
